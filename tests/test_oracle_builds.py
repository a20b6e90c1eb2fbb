"""The vectorized timing builds of the oracle (liborc_v3.so, liborc_v4.so),
which the CPU baseline of bench.py times, give the bits of the portable
build the tests check against: the same Born forward and adjoint, compared
in two fresh processes."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_SCRIPT = r"""
import sys
sys.path[:0] = [{root!r}, {tests!r}]
import numpy as np
from cases import make_case
from oracle import oracle as O
if {timing}:
    O.use_timing_build()
c = make_case(seed=3, nShot=1, sub=2, dts=0.002, nts=40)
wav = c.wavelet[None, :].copy()
m = np.random.default_rng(1).standard_normal((c.nx, c.nz)).astype(np.float32)
d = O.born_forward(c.grid, c.vel, c.oracle_sources(), [wav], c.oracle_receivers(), m, "f")
a = O.born_adjoint(c.grid, c.vel, c.oracle_sources(), [wav], c.oracle_receivers(), d, "f")
np.savez({out!r}, d=d, a=a, build=O._LIBNAME)
"""


def _run(tmp_path, timing):
    out = str(tmp_path / f"r{int(timing)}.npz")
    code = _SCRIPT.format(root=ROOT, tests=os.path.join(ROOT, "tests"), timing=timing, out=out)
    subprocess.run([sys.executable, "-c", code], check=True, env=dict(os.environ, OMP_NUM_THREADS="2"))
    return np.load(out)


def test_timing_build_matches_portable_build(tmp_path):
    base, fast = _run(tmp_path, False), _run(tmp_path, True)
    print("timing build:", str(fast["build"]))
    assert str(base["build"]) == "liborc.so"
    np.testing.assert_array_equal(base["d"], fast["d"])
    np.testing.assert_array_equal(base["a"], fast["a"])

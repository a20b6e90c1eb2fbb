"""Offset-extended Born on the GPU with the imaging and the secondary
sources batched over `extBatch` coarse samples: every batch size gives the
float oracle's bits, on the window-capture path (sub = 2) and the
accumulate path (sub = 1, 3), for offsets inside the register-tiled range
(hExt <= 20) and beyond it (hExt = 21, the untiled kernels)."""
import numpy as np
import pytest

from cases import make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _engine(c):
    from acoustic_isotropic_operators_amd import _capi as A
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=c.pad,
                  zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0, nExt=c.nExt,
                  extension=A.EXT_OFFSET)
    ctx = A.Context(params)
    ctx.set_velocity(c.vel)
    srcs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.src_args]
    recs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.rec_args]
    return ctx, srcs, recs


# nExt 21 and 31 image in offset groups of 7 and 14 (31: three groups, the
# last one partial); 43 takes the untiled kernels.
@pytest.mark.parametrize("sub,nExt", [(2, 21), (3, 11), (2, 43), (1, 3), (2, 31)])
def test_batched_offset_imaging_matches_oracle(sub, nExt):
    # 23 coarse samples: batches of 5 and 8 leave a partial last batch
    c = make_case(seed=61, nShot=3, nts=23, sub=sub, dts=0.001 * sub, nExt=nExt)
    ctx, srcs, recs = _engine(c)
    g, os_, or_ = c.grid, c.oracle_sources(), c.oracle_receivers()
    wav = c.wavelet[None, :].astype(np.float32)
    rng = np.random.default_rng(62)
    m = rng.standard_normal((nExt, c.nx, c.nz)).astype(np.float32)
    dr = rng.standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)
    of = O.bornext_forward(g, c.vel, os_, [wav], or_, m, "f")
    af = O.bornext_adjoint(g, c.vel, os_, [wav], or_, dr, "f")
    assert np.abs(of).max() > 0 and np.abs(af).max() > 0
    print(f"sub {sub} nExt {nExt}: |d| {np.abs(of).max():.3e} |m| {np.abs(af).max():.3e}", flush=True)
    for k in (1, 5, 8, 16):
        ctx.set_option("extBatch", k)
        d = np.zeros_like(dr)
        ctx.bornext_forward(m, d, srcs, wav, recs)
        np.testing.assert_array_equal(d, of, err_msg=f"forward, extBatch {k}")
        ma = np.zeros_like(m)
        ctx.bornext_adjoint(ma, dr, srcs, wav, recs)
        np.testing.assert_array_equal(ma, af, err_msg=f"adjoint, extBatch {k}")

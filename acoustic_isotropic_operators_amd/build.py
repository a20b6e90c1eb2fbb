"""Build the native parts of the package in-tree.

* ``libaiso.so``: the HIP engine (kernels.hip, engine.cpp, host.cpp) for gfx950,
  exporting the C-ABI declared in ``include/aiso.h``;
* the pybind11 modules ``pyAcoustic_iso_float_{nl,born,born_ext,tomo,wemva}`` that mirror the
  reference's module names on top of that C-ABI.

Run ``python -m acoustic_isotropic_operators_amd.build`` (or ``build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
ARCH = os.environ.get("AISO_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # keep hipcc from emitting packed FP32 instructions
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # no fused multiply-adds: every kernel variant then rounds every
             # product and sum the same way, as the reference-order CPU
             # restatement does, so results do not depend on the variant
             "-ffp-contract=off",
             "-I" + INC, "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({r.returncode})")
    return r


RESOURCE_USAGE = os.path.join(PKG, "_build", "kernels_resource_usage.txt")


def build_lib():
    from concurrent.futures import ThreadPoolExecutor
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    cmds, objs = [], []
    for src in ("kernels.hip", "noalloc.hip", "engine.cpp", "passes.cpp", "windows.cpp", "host.cpp"):
        obj = os.path.join(bdir, src.replace(".", "_") + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        # the kernels' register, scratch and occupancy figures, kept for the tests
        remarks = ["-Rpass-analysis=kernel-resource-usage"] if src == "kernels.hip" else []
        cmds.append([HIPCC, *HIP_FLAGS, *remarks, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # one compiler per object, side by side (each is single-threaded)
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(1, (os.cpu_count() or 1) // 2))) as ex:
        results = list(ex.map(_run, cmds))
    with open(RESOURCE_USAGE, "w") as f:
        f.write(results[0].stderr)
    out = os.path.join(PKG, "libaiso.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs])
    return out


def build_pybind():
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    inc = ["-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"], "-I" + INC, "-I" + CSRC]
    outs = []
    for name in ("pyAcoustic_iso_float_nl", "pyAcoustic_iso_float_born", "pyAcoustic_iso_float_born_ext",
                 "pyAcoustic_iso_float_tomo", "pyAcoustic_iso_float_wemva"):
        src = os.path.join(CSRC, "pybind", name + ".cpp")
        out = os.path.join(PKG, name + ext)
        _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", *inc, src,
              "-L" + PKG, "-laiso", "-Wl,-rpath,$ORIGIN", "-o", out])
        outs.append(out)
    return outs


def build():
    lib = build_lib()
    mods = build_pybind()
    return [lib, *mods]


if __name__ == "__main__":
    for p in build():
        print(p)

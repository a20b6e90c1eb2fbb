"""CPU oracle for the 2D acoustic isotropic FD engine -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker or the CPU
baseline.  The product package never imports it.

The arithmetic lives in ``liborc.so`` (C, ``aiso_oracle_body.h``), built by
``oracle/Makefile``, in float (``prec='f'``, follows the reference's
acoustic_iso_float) and double (``prec='d'``, follows acoustic_iso_double, the
FP64 oracle).  This module only sequences the per-shot C calls the way the
reference's host classes do:

* ``nl_forward`` / ``nl_adjoint``: F/src/nonlinearPropGpu.cpp:119-211 inside the
  shot loop of F/src/nonlinearPropShotsGpu.cpp:62-152,257-359;
* ``born_forward`` / ``born_adjoint``: F/src/BornGpu.cpp:29-89 with the source
  chain of F/include/seismicOperator2D.cpp:151-180 inside
  F/src/BornShotsGpu.cpp:60-143,245-339;
* ``bornext_forward`` / ``bornext_adjoint``: F/src/BornExtGpu.cpp:33-134 inside
  F/src/BornExtShotsGpu.cpp (same shot loop);
* ``borntime_forward`` / ``borntime_adjoint``: the same with extension=time
  (F/src/BornExtShotsGpuFunctions.cu:269-418, 1525-1678);
* ``tomo_forward`` / ``tomo_adjoint``: F/src/tomoExtGpu.cpp:34-100 inside
  F/src/tomoExtShotsGpu.cpp:62-145 (same shot loop), extension=offset or
  time;
* ``wemva_forward`` / ``wemva_adjoint``: F/src/wemvaExtGpu.cpp inside
  F/src/wemvaExtShotsGpu.cpp (same shot loop, observed data per shot),
  extension=offset or time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
FAT = 5


_LIBNAME = "liborc.so"


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, _LIBNAME)
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        _LIB = C.CDLL(path)
    return _LIB


def _cpu_flags():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def use_timing_build():
    """Switch to the vectorized build this CPU supports (liborc_v4.so with
    AVX-512, liborc_v3.so with AVX2/FMA, else the portable liborc.so), for
    the CPU baseline's timing; same bits as the portable build.  Call before
    the first oracle call.  Returns the build's file name."""
    global _LIB, _LIBNAME
    flags = _cpu_flags()
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    v3 = {"avx", "avx2", "fma", "bmi1", "bmi2", "movbe", "f16c"}
    name = "liborc.so"
    if v4 <= flags and os.path.exists(os.path.join(_HERE, "liborc_v4.so")):
        name = "liborc_v4.so"
    elif v3 <= flags and os.path.exists(os.path.join(_HERE, "liborc_v3.so")):
        name = "liborc_v3.so"
    if name != _LIBNAME:
        _LIBNAME, _LIB = name, None
    return name


class _Grid(C.Structure):
    _fields_ = [("nz", C.c_int), ("nx", C.c_int), ("nts", C.c_int), ("sub", C.c_int),
                ("minPad", C.c_int), ("nExt", C.c_int), ("dz", C.c_double),
                ("dx", C.c_double), ("dts", C.c_double), ("ots", C.c_double),
                ("alphaCos", C.c_double), ("freeSurface", C.c_int)]


@dataclass
class Grid:
    """Padded FD grid (sizes include the FAT ring), as fdParam holds it."""
    nz: int
    nx: int
    nts: int
    sub: int
    minPad: int
    dz: float
    dx: float
    dts: float
    ots: float = 0.0
    alphaCos: float = 0.99
    nExt: int = 1
    freeSurface: int = 0

    @property
    def ntw(self):
        return (self.nts - 1) * self.sub + 1

    def c(self):
        return _Grid(self.nz, self.nx, self.nts, self.sub, self.minPad, self.nExt,
                     self.dz, self.dx, self.dts, self.ots, self.alphaCos, self.freeSurface)


def _dt(prec):
    return np.float32 if prec == "f" else np.float64


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _fn(name, prec):
    return getattr(_lib(), f"{name}_{prec}")


def tables(g: Grid, prec="f"):
    """Laplacian coefficients, interpolation filter and sponge table."""
    dt = _dt(prec)
    zc, xc = np.zeros(6, dt), np.zeros(6, dt)
    f = np.zeros(2 * (g.sub + 1), dt)
    cd = np.zeros(max(g.minPad, 1), dt)
    gc = g.c()
    _fn("orc_tables", prec)(C.byref(gc), _p(zc), _p(xc), _p(f), _p(cd))
    return zc, xc, f, cd[:g.minPad]


def vel_scales(g: Grid, vel, prec="f"):
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dtype=dt)
    v2, rf = np.zeros_like(vel), np.zeros_like(vel)
    gc = g.c()
    _fn("orc_vel_scales", prec)(C.byref(gc), _p(vel), _p(v2), _p(rf))
    return v2, rf


def step_fwd(g, zc, xc, o, c, n, v2dt2, prec="f"):
    gc = g.c()
    _fn("orc_step_fwd", prec)(C.byref(gc), _p(zc), _p(xc), _p(o), _p(c), _p(n), _p(v2dt2))


def step_adj(g, zc, xc, o, c, n, v2dt2, prec="f"):
    gc = g.c()
    tmp = np.zeros_like(c)
    _fn("orc_step_adj", prec)(C.byref(gc), _p(zc), _p(xc), _p(o), _p(c), _p(n), _p(v2dt2), _p(tmp))


def damp(g, cd, p1, p2, prec="f"):
    gc = g.c()
    _fn("orc_damp", prec)(C.byref(gc), _p(cd), _p(p1), _p(p2))


def pad_plus(n, pad, block=16):
    return _lib().orc_pad_plus(int(n), int(pad), int(block))


# ----------------------------------------------------------------- geometry --
class Device:
    """deviceGpu restatement: raw corner list, weights, unique regular points."""

    def __init__(self, gridIdx, weight, nbCorner, nt):
        self.gridIdx = np.ascontiguousarray(gridIdx, dtype=np.int32)
        self.weight = weight
        self.nbCorner = nbCorner
        self.nDev = len(self.gridIdx) // nbCorner
        self.nt = nt
        reg = np.zeros(len(self.gridIdx), np.int32)
        cr = np.zeros(len(self.gridIdx), np.int32)
        n = _lib().orc_unique(len(self.gridIdx), _p(self.gridIdx), _p(reg), _p(cr))
        self.regPos = reg[:n].copy()
        self.cornerReg = cr
        self.nReg = n

    @classmethod
    def grid(cls, nzDev, ozDev, dzDev, nxDev, oxDev, dxDev, nzVel, nxVel, nt,
             dipole=0, zShift=0, xShift=0):
        """Constructor #3 (regular device grid in FD-grid indices)."""
        nb = 8 if dipole == 1 else 4
        n = nzDev * nxDev
        gi = np.zeros(n * nb, np.int32)
        w = np.zeros(n * nb, np.float32)
        r = _lib().orc_dev_grid(nzDev, ozDev, dzDev, nxDev, oxDev, dxDev, nzVel, nxVel,
                                dipole, zShift, xShift, _p(gi), _p(w))
        if r < 0:
            raise RuntimeError("device out of bounds")
        return cls(gi, w, nb, nt)

    @classmethod
    def ints(cls, zGrid, xGrid, nzVel, nxVel, nt, dipole=0, zShift=0, xShift=0):
        """Constructor #2 (explicit integer grid positions)."""
        nb = 8 if dipole == 1 else 4
        z = np.ascontiguousarray(zGrid, np.int32)
        x = np.ascontiguousarray(xGrid, np.int32)
        gi = np.zeros(len(z) * nb, np.int32)
        w = np.zeros(len(z) * nb, np.float32)
        r = _lib().orc_dev_ints(len(z), _p(z), _p(x), nzVel, nxVel, dipole, zShift, xShift,
                                _p(gi), _p(w))
        if r < 0:
            raise RuntimeError("device out of bounds")
        return cls(gi, w, nb, nt)

    @classmethod
    def coords(cls, zCoord, xCoord, nzVel, nxVel, oz, dz, ox, dx, nt, dipole=0,
               zShift=0.0, xShift=0.0, prec="f"):
        """Constructor #1 (physical coordinates, bilinear weights)."""
        dt = _dt(prec)
        z = np.ascontiguousarray(zCoord, dt)
        x = np.ascontiguousarray(xCoord, dt)
        nb = 8 if dipole == 1 else 4
        gi = np.zeros(len(z) * nb, np.int32)
        w = np.zeros(len(z) * nb, dt)
        f = _fn("orc_dev_coords", prec)
        f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double,
                      C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                      C.c_void_p, C.c_void_p]
        r = f(len(z), _p(z), _p(x), nzVel, nxVel, oz, dz, ox, dx, dipole, zShift, xShift,
              _p(gi), _p(w))
        if r < 0:
            raise RuntimeError("device out of bounds")
        return cls(gi, w, nb, nt)

    def _w(self, prec):
        return np.ascontiguousarray(self.weight, _dt(prec))

    def forward(self, mReg, add=False, out=None, prec="f"):
        """Regular traces [nReg][nt] -> device traces [nDev][nt]."""
        dt = _dt(prec)
        mReg = np.ascontiguousarray(mReg, dt)
        nt = mReg.shape[-1]
        out = np.zeros((self.nDev, nt), dt) if out is None else out
        w = self._w(prec)
        _fn("orc_dev_fwd", prec)(self.nDev, self.nbCorner, _p(self.cornerReg), _p(w), nt,
                                 int(add), _p(mReg), _p(out))
        return out

    def adjoint(self, dIrreg, add=False, out=None, prec="f"):
        """Device traces [nDev][nt] -> regular traces [nReg][nt]."""
        dt = _dt(prec)
        dIrreg = np.ascontiguousarray(dIrreg, dt)
        nt = dIrreg.shape[-1]
        out = np.zeros((self.nReg, nt), dt) if out is None else out
        w = self._w(prec)
        _fn("orc_dev_adj", prec)(self.nDev, self.nbCorner, _p(self.cornerReg), _p(w), nt,
                                 int(add), _p(out), _p(dIrreg))
        return out


# ------------------------------------------------------ host signal chain --
def interp_fwd(g: Grid, m, add=False, out=None, prec="f"):
    dt = _dt(prec)
    m = np.ascontiguousarray(m, dt)
    nDev = m.shape[0]
    out = np.zeros((nDev, g.ntw), dt) if out is None else out
    f = _fn("orc_interp_fwd", prec)
    f.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    f(g.nts, g.dts, g.ots, g.sub, nDev, int(add), _p(m), _p(out))
    return out


def interp_adj(g: Grid, d, add=False, out=None, prec="f"):
    dt = _dt(prec)
    d = np.ascontiguousarray(d, dt)
    nDev = d.shape[0]
    out = np.zeros((nDev, g.nts), dt) if out is None else out
    f = _fn("orc_interp_adj", prec)
    f.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    f(g.nts, g.dts, g.ots, g.sub, nDev, int(add), _p(out), _p(d))
    return out


def d2t_fwd(nt, dt_, m, add=False, out=None, prec="f"):
    dt = _dt(prec)
    m = np.ascontiguousarray(m, dt)
    out = np.zeros_like(m) if out is None else out
    f = _fn("orc_d2t_fwd", prec)
    f.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    f(nt, dt_, m.shape[0], int(add), _p(m), _p(out))
    return out


def scale_source(g: Grid, vel, pos, sig, prec="f"):
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    pos = np.ascontiguousarray(pos, np.int32)
    gc = g.c()
    _fn("orc_scale_source", prec)(C.byref(gc), _p(vel), _p(pos), len(pos), sig.shape[-1], _p(sig))
    return sig


# ---------------------------------------------------- per-shot propagators --
def nl_fwd_reg(g, v2dt2, sig, srcPos, recPos, prec="f"):
    dreg = np.zeros((len(recPos), g.nts), _dt(prec))
    gc = g.c()
    _fn("orc_nl_fwd", prec)(C.byref(gc), _p(v2dt2), _p(sig), _p(srcPos), len(srcPos),
                            _p(recPos), len(recPos), _p(dreg))
    return dreg


def nl_adj_reg(g, v2dt2, dreg, srcPos, recPos, prec="f"):
    sig = np.zeros((len(srcPos), g.ntw), _dt(prec))
    gc = g.c()
    _fn("orc_nl_adj", prec)(C.byref(gc), _p(v2dt2), _p(sig), _p(srcPos), len(srcPos),
                            _p(recPos), len(recPos), _p(np.ascontiguousarray(dreg)))
    return sig


def nl_vgrad_reg(g, v2dt2, sig, srcPos, recPos, dres, prec="f"):
    """One shot: wave-equation part of dJ/dv2dt2 and the adjoint recorded at
    the source points (orc_nl_vgrad)."""
    dt = _dt(prec)
    grad = np.zeros((g.nx, g.nz), dt)
    sigAdj = np.zeros((len(srcPos), g.ntw), dt)
    gc = g.c()
    _fn("orc_nl_vgrad", prec)(C.byref(gc), _p(v2dt2), _p(sig), _p(srcPos), len(srcPos),
                              _p(recPos), len(recPos), _p(np.ascontiguousarray(dres, dt)), _p(grad), _p(sigAdj))
    return grad, sigAdj


def _born_call(name, g, v2dt2, refl, sig, srcPos, recPos, model, dreg, prec):
    W = np.zeros((g.nts, g.nx, g.nz), _dt(prec))
    gc = g.c()
    _fn(name, prec)(C.byref(gc), _p(v2dt2), _p(refl), _p(sig), _p(srcPos), len(srcPos),
                    _p(recPos), len(recPos), _p(model), _p(dreg), _p(W))


def born_fwd_reg(g, v2dt2, refl, sig, srcPos, recPos, model, prec="f", ext=False):
    dreg = np.zeros((len(recPos), g.nts), _dt(prec))
    model = np.ascontiguousarray(model, _dt(prec))
    _born_call("orc_bornext_fwd" if ext else "orc_born_fwd", g, v2dt2, refl, sig, srcPos,
               recPos, model, dreg, prec)
    return dreg


def born_adj_reg(g, v2dt2, refl, sig, srcPos, recPos, dreg, prec="f", ext=False):
    shape = (g.nExt, g.nx, g.nz) if ext else (g.nx, g.nz)
    model = np.zeros(shape, _dt(prec))
    _born_call("orc_bornext_adj" if ext else "orc_born_adj", g, v2dt2, refl, sig, srcPos,
               recPos, model, np.ascontiguousarray(dreg, _dt(prec)), prec)
    return model


def borntime_fwd_reg(g, v2dt2, refl, sig, srcPos, recPos, model, sloth=0, prec="f"):
    dt = _dt(prec)
    dreg = np.zeros((len(recPos), g.nts), dt)
    W = np.zeros((g.nts, g.nx, g.nz), dt)
    gc = g.c()
    _fn("orc_borntime_fwd", prec)(C.byref(gc), _p(v2dt2), _p(refl), int(sloth), _p(sig), _p(srcPos), len(srcPos),
                                  _p(recPos), len(recPos), _p(np.ascontiguousarray(model, dt)), _p(dreg), _p(W))
    return dreg


def borntime_adj_reg(g, v2dt2, refl, sig, srcPos, recPos, dreg, sloth=0, prec="f"):
    dt = _dt(prec)
    model = np.zeros((g.nExt, g.nx, g.nz), dt)
    W = np.zeros((g.nts, g.nx, g.nz), dt)
    gc = g.c()
    _fn("orc_borntime_adj", prec)(C.byref(gc), _p(v2dt2), _p(refl), int(sloth), _p(sig), _p(srcPos), len(srcPos),
                                  _p(recPos), len(recPos), _p(model), _p(np.ascontiguousarray(dreg, dt)), _p(W))
    return model


def _tomo_fns(name, prec):
    f = _fn(name, prec)
    f.restype = None
    return f


def tomo_fwd_reg(g, v2dt2, refl, sig, srcPos, recPos, model, ext_refl, leg1=1, leg2=1, prec="f", time=False,
                 sloth=0):
    dt = _dt(prec)
    dreg = np.zeros((len(recPos), g.nts), dt)
    W = np.zeros((g.nts, g.nx, g.nz), dt)
    P = np.zeros_like(W)
    gc = g.c()
    _tomo_fns("orc_tomo_fwd", prec)(C.byref(gc), _p(v2dt2), _p(refl), _p(sig), _p(srcPos), len(srcPos),
                                    _p(recPos), len(recPos), _p(np.ascontiguousarray(model, dt)),
                                    _p(np.ascontiguousarray(ext_refl, dt)), int(leg1), int(leg2), int(time),
                                    int(sloth), _p(dreg), _p(W), _p(P))
    return dreg


def tomo_adj_reg(g, v2dt2, refl, sig, srcPos, recPos, dreg, ext_refl, leg1=1, leg2=1, prec="f", time=False,
                 sloth=0):
    dt = _dt(prec)
    model = np.zeros((g.nx, g.nz), dt)
    W = np.zeros((g.nts, g.nx, g.nz), dt)
    P = np.zeros_like(W)
    gc = g.c()
    _tomo_fns("orc_tomo_adj", prec)(C.byref(gc), _p(v2dt2), _p(refl), _p(sig), _p(srcPos), len(srcPos),
                                    _p(recPos), len(recPos), _p(model), _p(np.ascontiguousarray(dreg, dt)),
                                    _p(np.ascontiguousarray(ext_refl, dt)), int(leg1), int(leg2), int(time),
                                    int(sloth), _p(W), _p(P))
    return model


def wemva_fwd_reg(g, v2dt2, refl, sig, srcPos, recPos, model, dobs, leg1=1, leg2=1, prec="f", time=False):
    dt = _dt(prec)
    image = np.zeros((g.nExt, g.nx, g.nz), dt)
    W = np.zeros((g.nts, g.nx, g.nz), dt)
    P = np.zeros_like(W)
    gc = g.c()
    _tomo_fns("orc_wemva_fwd", prec)(C.byref(gc), _p(v2dt2), _p(refl), _p(sig), _p(srcPos), len(srcPos),
                                     _p(recPos), len(recPos), _p(np.ascontiguousarray(model, dt)),
                                     _p(np.ascontiguousarray(dobs, dt)), int(leg1), int(leg2), int(time),
                                     _p(image), _p(W), _p(P))
    return image


def wemva_adj_reg(g, v2dt2, refl, sig, srcPos, recPos, image, dobs, leg1=1, leg2=1, prec="f", time=False):
    dt = _dt(prec)
    model = np.zeros((g.nx, g.nz), dt)
    W = np.zeros((g.nts, g.nx, g.nz), dt)
    P = np.zeros_like(W)
    gc = g.c()
    _tomo_fns("orc_wemva_adj", prec)(C.byref(gc), _p(v2dt2), _p(refl), _p(sig), _p(srcPos), len(srcPos),
                                     _p(recPos), len(recPos), _p(model), _p(np.ascontiguousarray(image, dt)),
                                     _p(np.ascontiguousarray(dobs, dt)), int(leg1), int(leg2), int(time),
                                     _p(W), _p(P))
    return model


def dt2_coarse(g, W, prec="f"):
    """Second time derivative of a coarse wavefield [nts][nx][nz] (interior
    points), zero beyond both ends, as the tomographic operator applies it."""
    W = np.array(W, _dt(prec), copy=True, order="C")
    gc = g.c()
    _tomo_fns("orc_dt2_coarse", prec)(C.byref(gc), _p(W))
    return W


# ------------------------------------------------- multi-shot operators --
def _shot_geom(lst, i):
    return lst[0] if len(lst) == 1 else lst[i]


def born_source_fine(g, vel, src: Device, wavelet, prec="f"):
    """S = I_t(D_v(D2_t(R_src^T w))) -- F/include/seismicOperator2D.cpp:151-180."""
    reg = src.adjoint(wavelet, prec=prec)
    reg2 = d2t_fwd(g.nts, g.dts, reg, prec=prec)
    scale_source(g, vel, src.regPos, reg2, prec=prec)
    return interp_fwd(g, reg2, prec=prec)


def nl_forward(g, vel, sources, receivers, model, prec="f"):
    """model [nShotSig][nSrcDev][nts] -> data [nShot][nRec][nts]."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, _ = vel_scales(g, vel, prec)
    nShot = len(sources)
    nRec = _shot_geom(receivers, 0).nDev
    data = np.zeros((nShot, nRec, g.nts), dt)
    for s in range(nShot):
        src, rec = sources[s], _shot_geom(receivers, s)
        m = model[0] if model.shape[0] == 1 else model[s]
        reg = src.adjoint(m, prec=prec)
        scale_source(g, vel, src.regPos, reg, prec=prec)
        fine = interp_fwd(g, reg, prec=prec)
        dreg = nl_fwd_reg(g, v2, fine, src.regPos, rec.regPos, prec)
        sl = rec.forward(dreg, prec=prec)
        data[s] += sl
    return data


def nl_velocity_gradient(g, vel, sources, receivers, model, dgrad, prec="f"):
    """dJ/dvel of a misfit J whose derivative with respect to the data of
    nl_forward(vel) is dgrad [nShot][nRec][nts] (for J = |d - d_obs|^2 / 2,
    dgrad = d - d_obs).  Procedure (the engine follows it step for step):
    per shot the wave-equation part (orc_nl_vgrad) summed in shot order; then,
    shot by shot and source point by source point, the source-scaling part
    sum_t sigAdj * u (u: the unscaled fine source signal, in double, rounded
    once); finally the chain factor d(v^2 dtw^2)/dv = 2 v dtw^2."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, _ = vel_scales(g, vel, prec)
    G2 = np.zeros((g.nx, g.nz), dt)
    terms = []
    for s in range(len(sources)):
        src, rec = sources[s], _shot_geom(receivers, s)
        m = model[0] if model.shape[0] == 1 else model[s]
        reg = src.adjoint(m, prec=prec)
        u = interp_fwd(g, reg.copy(), prec=prec)
        scale_source(g, vel, src.regPos, reg, prec=prec)
        fine = interp_fwd(g, reg, prec=prec)
        dres = rec.adjoint(np.ascontiguousarray(dgrad[s], dt), prec=prec)
        gs, sigAdj = nl_vgrad_reg(g, v2, fine, src.regPos, rec.regPos, dres, prec)
        G2 += gs
        terms.append([(int(pos), float(np.dot(sigAdj[r].astype(np.float64), u[r].astype(np.float64))))
                      for r, pos in enumerate(src.regPos)])
    flat = G2.reshape(-1)
    for shot in terms:
        for pos, val in shot:
            flat[pos] += dt(val)
    dtw = dt(g.dts) / dt(g.sub)
    return G2 * ((dt(2.0) * vel) * dtw * dtw)


def nl_adjoint(g, vel, sources, receivers, data, nShotSig=1, prec="f"):
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, _ = vel_scales(g, vel, prec)
    nShot = len(sources)
    nSrcDev = sources[0].nDev
    model = np.zeros((nShotSig, nSrcDev, g.nts), dt)
    acc = np.zeros((nSrcDev, g.nts), dt)
    for s in range(nShot):
        src, rec = sources[s], _shot_geom(receivers, s)
        dreg = rec.adjoint(data[s], prec=prec)
        fine = nl_adj_reg(g, v2, dreg, src.regPos, rec.regPos, prec)
        reg = interp_adj(g, fine, prec=prec)
        scale_source(g, vel, src.regPos, reg, prec=prec)
        if nShotSig == 1:
            src.forward(reg, add=True, out=acc, prec=prec)
        else:
            sl = src.forward(reg, prec=prec)
            model[s] += sl
    if nShotSig == 1:
        model[0] += acc
    return model


def born_forward(g, vel, sources, wavelets, receivers, model, prec="f", ext=False):
    """model [nx][nz] (or [nExt][nx][nz]) -> data [nShot][nRec][nts]."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    nShot = len(sources)
    nRec = _shot_geom(receivers, 0).nDev
    data = np.zeros((nShot, nRec, g.nts), dt)
    for s in range(nShot):
        src, rec = sources[s], _shot_geom(receivers, s)
        fine = born_source_fine(g, vel, src, _shot_geom(wavelets, s), prec)
        dreg = born_fwd_reg(g, v2, rf, fine, src.regPos, rec.regPos, model, prec, ext)
        data[s] += rec.forward(dreg, prec=prec)
    return data


def born_adjoint(g, vel, sources, wavelets, receivers, data, prec="f", ext=False):
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    shape = (g.nExt, g.nx, g.nz) if ext else (g.nx, g.nz)
    image = np.zeros(shape, dt)
    for s in range(len(sources)):
        src, rec = sources[s], _shot_geom(receivers, s)
        fine = born_source_fine(g, vel, src, _shot_geom(wavelets, s), prec)
        dreg = rec.adjoint(data[s], prec=prec)
        image += born_adj_reg(g, v2, rf, fine, src.regPos, rec.regPos, dreg, prec, ext)
    return image


def nl_forward_wavefield(g, vel, sources, receivers, model, shot, prec="f"):
    """The coarse wavefield [nts][nx][nz] of shot `shot` of nl_forward, as the
    reference's forwardWavefield saves it (propShotsFwdGpuWavefield,
    F/src/nonlinearShotsGpuFunctions.cu:310-390)."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, _ = vel_scales(g, vel, prec)
    src, rec = sources[shot], _shot_geom(receivers, shot)
    m = model[0] if model.shape[0] == 1 else model[shot]
    reg = src.adjoint(m, prec=prec)
    scale_source(g, vel, src.regPos, reg, prec=prec)
    fine = interp_fwd(g, reg, prec=prec)
    dreg = np.zeros((len(rec.regPos), g.nts), dt)
    wf = np.zeros((g.nts, g.nx, g.nz), dt)
    gc = g.c()
    _fn("orc_nl_fwd_wfld", prec)(C.byref(gc), _p(v2), _p(fine), _p(src.regPos), len(src.regPos),
                                 _p(rec.regPos), len(rec.regPos), _p(dreg), _p(wf))
    return wf


def nl_adjoint_wavefield(g, vel, sources, receivers, data, shot, prec="f"):
    """The coarse adjoint wavefield [nts][nx][nz] of shot `shot` of
    nl_adjoint (propShotsAdjGpuWavefield, F/src/nonlinearShotsGpuFunctions.cu:653-740)."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, _ = vel_scales(g, vel, prec)
    src, rec = sources[shot], _shot_geom(receivers, shot)
    dreg = np.ascontiguousarray(rec.adjoint(data[shot], prec=prec), dt)
    sig = np.zeros((len(src.regPos), g.ntw), dt)
    wf = np.zeros((g.nts, g.nx, g.nz), dt)
    gc = g.c()
    _fn("orc_nl_adj_wfld", prec)(C.byref(gc), _p(v2), _p(sig), _p(src.regPos), len(src.regPos),
                                 _p(rec.regPos), len(rec.regPos), _p(dreg), _p(wf))
    return wf


def born_wavefields(g, vel, sources, wavelets, receivers, vec, shot, adjoint, prec="f"):
    """The source and secondary wavefields [nts][nx][nz] of shot `shot` that
    the reference's forwardWavefield / adjointWavefield save
    (BornShotsFwdGpuWavefield / BornShotsAdjGpuWavefield,
    F/src/BornShotsGpuFunctions.cu:390-520, 959-1080): the coarse source
    wavefield, and the scattered wavefield (forward; vec = model) or the
    receiver wavefield scaled by 2/v^3 v^2 dtw^2 (adjoint; vec = data)."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    src, rec = sources[shot], _shot_geom(receivers, shot)
    fine = born_source_fine(g, vel, src, _shot_geom(wavelets, shot), prec)
    W = np.zeros((g.nts, g.nx, g.nz), dt)
    Wsec = np.zeros((g.nts, g.nx, g.nz), dt)
    gc = g.c()
    if adjoint:
        dreg = np.ascontiguousarray(rec.adjoint(vec[shot], prec=prec), dt)
        model = np.zeros((g.nx, g.nz), dt)
        _fn("orc_born_adj_wfld", prec)(C.byref(gc), _p(v2), _p(rf), _p(fine), _p(src.regPos), len(src.regPos),
                                       _p(rec.regPos), len(rec.regPos), _p(model), _p(dreg), _p(W), _p(Wsec))
    else:
        dreg = np.zeros((len(rec.regPos), g.nts), dt)
        model = np.ascontiguousarray(vec, dt)
        _fn("orc_born_fwd_wfld", prec)(C.byref(gc), _p(v2), _p(rf), _p(fine), _p(src.regPos), len(src.regPos),
                                       _p(rec.regPos), len(rec.regPos), _p(model), _p(dreg), _p(W), _p(Wsec))
    return W, Wsec


def bornext_forward(g, vel, sources, wavelets, receivers, model, prec="f"):
    return born_forward(g, vel, sources, wavelets, receivers, model, prec, ext=True)


def bornext_adjoint(g, vel, sources, wavelets, receivers, data, prec="f"):
    return born_adjoint(g, vel, sources, wavelets, receivers, data, prec, ext=True)


def borntime_forward(g, vel, sources, wavelets, receivers, model, sloth=0, prec="f"):
    """Time-lag extended Born (extension=time): model [nExt][nx][nz] -> data."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    nShot = len(sources)
    nRec = _shot_geom(receivers, 0).nDev
    data = np.zeros((nShot, nRec, g.nts), dt)
    for s in range(nShot):
        src, rec = sources[s], _shot_geom(receivers, s)
        fine = born_source_fine(g, vel, src, _shot_geom(wavelets, s), prec)
        dreg = borntime_fwd_reg(g, v2, rf, fine, src.regPos, rec.regPos, model, sloth, prec)
        data[s] += rec.forward(dreg, prec=prec)
    return data


def borntime_adjoint(g, vel, sources, wavelets, receivers, data, sloth=0, prec="f"):
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    image = np.zeros((g.nExt, g.nx, g.nz), dt)
    for s in range(len(sources)):
        src, rec = sources[s], _shot_geom(receivers, s)
        fine = born_source_fine(g, vel, src, _shot_geom(wavelets, s), prec)
        dreg = rec.adjoint(data[s], prec=prec)
        image += borntime_adj_reg(g, v2, rf, fine, src.regPos, rec.regPos, dreg, sloth, prec)
    return image


def tomo_forward(g, vel, sources, wavelets, receivers, model, ext_refl, leg1=1, leg2=1, prec="f", time=False,
                 sloth=0):
    """Extended tomographic operator (tomoExtShotsGpu, extension=offset, or
    extension=time with ``time=True``): velocity perturbation [nx][nz] ->
    data [nShot][nRec][nts], about the extended reflectivity ext_refl
    [nExt][nx][nz]."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    nShot = len(sources)
    nRec = _shot_geom(receivers, 0).nDev
    data = np.zeros((nShot, nRec, g.nts), dt)
    for s in range(nShot):
        src, rec = sources[s], _shot_geom(receivers, s)
        fine = born_source_fine(g, vel, src, _shot_geom(wavelets, s), prec)
        dreg = tomo_fwd_reg(g, v2, rf, fine, src.regPos, rec.regPos, model, ext_refl, leg1, leg2, prec, time, sloth)
        data[s] += rec.forward(dreg, prec=prec)
    return data


def wemva_forward(g, vel, sources, wavelets, receivers, dobs, model, leg1=1, leg2=1, prec="f", time=False):
    """Extended WEMVA operator (wemvaExtShotsGpu, extension=offset, or
    extension=time with ``time=True``): velocity perturbation [nx][nz] ->
    extended image [nExt][nx][nz], about the receiver wavefields of the
    observed data dobs [nShot][nRec][nts].  Like the reference, it ignores a
    free surface."""
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    image = np.zeros((g.nExt, g.nx, g.nz), dt)
    for s in range(len(sources)):
        src, rec = sources[s], _shot_geom(receivers, s)
        fine = born_source_fine(g, vel, src, _shot_geom(wavelets, s), prec)
        dreg = rec.adjoint(dobs[s], prec=prec)
        image += wemva_fwd_reg(g, v2, rf, fine, src.regPos, rec.regPos, model, dreg, leg1, leg2, prec, time)
    return image


def wemva_adjoint(g, vel, sources, wavelets, receivers, dobs, image, leg1=1, leg2=1, prec="f", time=False):
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    model = np.zeros((g.nx, g.nz), dt)
    for s in range(len(sources)):
        src, rec = sources[s], _shot_geom(receivers, s)
        fine = born_source_fine(g, vel, src, _shot_geom(wavelets, s), prec)
        dreg = rec.adjoint(dobs[s], prec=prec)
        model += wemva_adj_reg(g, v2, rf, fine, src.regPos, rec.regPos, image, dreg, leg1, leg2, prec, time)
    return model


def tomo_adjoint(g, vel, sources, wavelets, receivers, data, ext_refl, leg1=1, leg2=1, prec="f", time=False,
                 sloth=0):
    dt = _dt(prec)
    vel = np.ascontiguousarray(vel, dt)
    v2, rf = vel_scales(g, vel, prec)
    model = np.zeros((g.nx, g.nz), dt)
    for s in range(len(sources)):
        src, rec = sources[s], _shot_geom(receivers, s)
        fine = born_source_fine(g, vel, src, _shot_geom(wavelets, s), prec)
        dreg = rec.adjoint(data[s], prec=prec)
        model += tomo_adj_reg(g, v2, rf, fine, src.regPos, rec.regPos, dreg, ext_refl, leg1, leg2, prec, time, sloth)
    return model

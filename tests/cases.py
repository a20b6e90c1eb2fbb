"""Seeded test cases shared by the oracle and GPU parity tests.

A case holds one padded FD grid, a smooth random velocity, per-shot source
geometries, one receiver geometry (or one per shot), a Ricker wavelet and
random model/data vectors, in the reference's layouts.  ``par`` is a
ParamObj carrying the reference's parfile keys.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from acoustic_isotropic_operators_amd.sep import ParamObj
from oracle import oracle as O

FAT = 5


def ricker(nts, dts, f0, t0):
    """Ricker wavelet (seis_utils/seis_utils_float/python/waveletFloatMain.py:95-99 formula)."""
    t = np.arange(nts) * dts - t0
    a = (np.pi * f0 * t) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


@dataclass
class Case:
    nz: int
    nx: int
    nts: int
    dts: float
    sub: int
    pad: int
    nShot: int
    nExt: int
    alpha: float
    vel: np.ndarray
    wavelet: np.ndarray
    src_args: list          # per shot: ("grid", args) | ("ints", (z, x)) | ("coords", (z, x))
    rec_args: list
    par: ParamObj
    dz: float = 10.0
    dx: float = 10.0
    extra: dict = field(default_factory=dict)

    @property
    def grid(self):
        return O.Grid(nz=self.nz, nx=self.nx, nts=self.nts, sub=self.sub, minPad=self.pad,
                      dz=self.dz, dx=self.dx, dts=self.dts, ots=0.0, alphaCos=self.alpha, nExt=self.nExt)

    # oracle geometries
    def _odev(self, spec):
        kind, args = spec
        if kind == "grid":
            return O.Device.grid(*args, self.nz, self.nx, self.nts)
        if kind == "ints":
            return O.Device.ints(args[0], args[1], self.nz, self.nx, self.nts)
        return O.Device.coords(np.asarray(args[0]), np.asarray(args[1]), self.nz, self.nx,
                               0.0, self.dz, 0.0, self.dx, self.nts)

    def oracle_sources(self):
        return [self._odev(s) for s in self.src_args]

    def oracle_receivers(self):
        return [self._odev(s) for s in self.rec_args]

    # product geometries (pybind deviceGpu)
    def _pdev(self, nl, spec):
        kind, args = spec
        v = self.vel_vector()
        if kind == "grid":
            return nl.deviceGpu(*args, v, self.nts, 0, 0, 0)
        if kind == "ints":
            return nl.deviceGpu(list(args[0]), list(args[1]), v, self.nts, 0, 0, 0)
        return nl.deviceGpu(np.asarray(args[0], np.float32), np.asarray(args[1], np.float32), v, self.nts, 0, 0.0, 0.0)

    def vel_vector(self):
        from acoustic_isotropic_operators_amd.sep import FloatVector, hypercube
        return FloatVector(hypercube(ns=[self.nz, self.nx], os_=[0.0, 0.0], ds=[self.dz, self.dx]), self.vel)

    def product_sources(self, nl):
        return [self._pdev(nl, s) for s in self.src_args]

    def product_receivers(self, nl):
        return [self._pdev(nl, s) for s in self.rec_args]


def make_case(seed=0, nz=58, nx=74, nts=60, dts=0.004, sub=4, pad=10, nShot=2, nExt=1,
              alpha=0.95, src_kind="grid", rec_kind="grid", per_shot_rec=False, src_iz=16,
              shot_dx=12):
    rng = np.random.default_rng(seed)
    z = np.linspace(0.0, 1.0, nz)[None, :]
    vel = 1800.0 + 700.0 * z + 60.0 * rng.standard_normal((nx, nz))
    k = np.ones(5) / 5.0
    for _ in range(3):  # smooth the perturbation
        vel = np.apply_along_axis(lambda r: np.convolve(r, k, mode="same"), 0, vel)
        vel = np.apply_along_axis(lambda r: np.convolve(r, k, mode="same"), 1, vel)
    vel = np.clip(vel, 1700.0, 2700.0).astype(np.float32)
    src = []
    for s in range(nShot):
        ix = 15 + shot_dx * s
        if src_kind == "grid":
            src.append(("grid", (1, src_iz, 1, 1, ix, 1)))
        elif src_kind == "ints":
            src.append(("ints", ([src_iz], [ix])))
        else:
            src.append(("coords", ([src_iz * 10.0 + 3.0], [ix * 10.0 + 4.0])))
    def rec_for(s):
        if rec_kind == "grid":
            return ("grid", (1, src_iz, 1, (nx - 2 * FAT - 12) // 2, FAT + 4, 2))
        if rec_kind == "coords":
            xs = np.arange(FAT + 4, nx - FAT - 6, 2.0) * 10.0 + 2.5 * (s + 1)
            return ("coords", (np.full(len(xs), src_iz * 10.0 + 4.0), xs))
        return ("ints", (list(range(FAT + 2, nz - FAT, 7)), [FAT + 8 + s] * len(range(FAT + 2, nz - FAT, 7))))
    rec = [rec_for(s) for s in range(nShot)] if per_shot_rec else [rec_for(0)]
    par = ParamObj([f"nts={nts}", f"dts={dts}", "ots=0.0", f"sub={sub}", f"nz={nz}", f"nx={nx}",
                    "dz=10.0", "dx=10.0", f"zPadMinus={pad}", f"zPadPlus={pad}", f"xPadMinus={pad}",
                    f"xPadPlus={pad}", "fat=5", "blockSize=16", f"alphaCos={alpha}", "fMax=20",
                    f"nShot={nShot}", "nGpu=1", f"nExt={nExt}",
                    f"extension={'offset' if nExt > 1 else 'none'}"])
    wav = ricker(nts, dts, 15.0, 0.05)
    return Case(nz, nx, nts, dts, sub, pad, nShot, nExt, alpha, vel, wav, src, rec, par)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))

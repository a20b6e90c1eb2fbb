"""Seeded synthetic workloads (velocity, reflectivity, wavelet, geometry).

* ``nonlinear_small``: one shot on a 200 x 400 grid (NL forward);
* ``born_lsrtm``: the Born forward+adjoint benchmark, 80 shots on 500 x 1200.

Grids are padded with the rule of F/src/padFileGpuMain.cpp:49-118 (pads
edge-replicated, FAT ring left at zero), devices are placed like
buildSourceGeometry/buildReceiversGeometry (PY/Acoustic_iso_float.py:237-456:
iz = zSource-1+zPadMinus+fat).  The data are synthetic, generated here.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

FAT = 5


def pad_plus(n, pad, block=16):
    """padPlus of F/src/padFileGpuMain.cpp:49-66 (ceil in float)."""
    total = 2 * pad + n
    return int(np.ceil(np.float32(total) / np.float32(block))) * block - n - pad


def pad_plus_free_surface(nz, zPad, block=16):
    """zPadPlus with a free surface, F/src/padFileGpuMain.cpp:37-47: no top
    padding, all of it below."""
    return int(np.ceil(np.float32(zPad + nz) / np.float32(block))) * block - nz


def pad_model(m, zPad, xPad, block=16, free_surface=False):
    """Edge-replicate a physical [nx][nz] model into the padded grid; FAT ring = 0.
    With a free surface the model starts on the first interior row."""
    nx, nz = m.shape
    zPlus, xPlus = pad_plus(nz, zPad, block), pad_plus(nx, xPad, block)
    if free_surface:
        zPlus, zPad = pad_plus_free_surface(nz, zPad, block), 0
    NZ, NX = nz + zPad + zPlus + 2 * FAT, nx + xPad + xPlus + 2 * FAT
    out = np.zeros((NX, NZ), np.float32)
    z0, x0 = FAT + zPad, FAT + xPad
    out[x0:x0 + nx, z0:z0 + nz] = m
    out[x0:x0 + nx, FAT:z0] = m[:, :1]
    out[x0:x0 + nx, z0 + nz:z0 + nz + zPlus] = m[:, -1:]
    out[FAT:x0, FAT:NZ - FAT] = out[x0:x0 + 1, FAT:NZ - FAT]
    out[x0 + nx:x0 + nx + xPlus, FAT:NZ - FAT] = out[x0 + nx - 1:x0 + nx, FAT:NZ - FAT]
    return out, zPlus, xPlus


def ricker(nts, dts, f0, t0):
    """Ricker wavelet, seis_utils/seis_utils_float/python/waveletFloatMain.py:95-99."""
    t = np.arange(nts, dtype=np.float64) * dts - t0
    a = (np.pi * f0 * t) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


@dataclass
class Workload:
    name: str
    params: dict          # aiso_fd_params fields
    vel: np.ndarray       # padded [NX][NZ]
    wavelet: np.ndarray   # [1][nts] (one point source per shot)
    src: list             # per shot (iz, ix) on the padded grid
    rec: tuple            # (nzDev, ozDev, dzDev, nxDev, oxDev, dxDev)
    model: np.ndarray     # padded reflectivity [NX][NZ]

    @property
    def nShot(self):
        return len(self.src)

    @property
    def interior(self):
        return (self.params["nz"] - 2 * FAT) * (self.params["nx"] - 2 * FAT)

    @property
    def fine_steps(self):
        return (self.params["nts"] - 1) * self.params["sub"]

    def parfile(self):
        """The reference's parfile keys for this workload."""
        p = self.params
        keys = dict(nts=p["nts"], dts=p["dts"], ots=0.0, sub=p["sub"], nz=p["nz"], nx=p["nx"], dz=p["dz"],
                    dx=p["dx"], zPadMinus=p["zPadMinus"], zPadPlus=p["zPadPlus"], xPadMinus=p["xPadMinus"],
                    xPadPlus=p["xPadPlus"], fat=5, blockSize=16, alphaCos=p["alphaCos"], fMax=p["fMax"],
                    nShot=self.nShot, nGpu=1)
        return [f"{k}={v}" for k, v in keys.items()]


def _smooth(field, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(field, sigma, mode="nearest")


def born_lsrtm(nShot=80, seed=2024, nts=2000, dts=0.002, sub=2, free_surface=False):
    """500 x 1200 grid at 10 m, 50-point sponge (alpha 0.95), 1500-4000 m/s
    gradient with +/-5 % smoothed random perturbations (sigma 300 m),
    nShot sources and 1200 receivers at 20 m depth, nts=2000 at 2 ms (sub 2,
    Courant ~0.42), 15 Hz Ricker, reflectivity = sparse N(0,1) spikes on 10
    dipping interfaces.  free_surface: the same with a free surface on the
    model's top row instead of the top sponge."""
    rng = np.random.default_rng(seed)
    nz, nx, d, pad = 500, 1200, 10.0, 50
    z = np.arange(nz) / (nz - 1)
    v = np.broadcast_to(1500.0 + 2500.0 * z, (nx, nz)).astype(np.float64)
    pert = _smooth(rng.standard_normal((nx, nz)), 30.0)
    pert /= np.abs(pert).max()
    v = (v * (1.0 + 0.05 * pert)).astype(np.float32)
    refl = np.zeros((nx, nz), np.float32)
    xs = np.arange(nx)
    for _ in range(10):
        z0 = rng.uniform(0.15, 0.95) * nz
        slope = rng.uniform(-0.15, 0.15)
        zi = np.clip(np.round(z0 + slope * (xs - nx / 2)).astype(int), 0, nz - 1)
        keep = rng.random(nx) < 0.5          # ~1 % of all model points carry a spike
        refl[xs[keep], zi[keep]] = rng.standard_normal(int(keep.sum())).astype(np.float32)
    vp, zPlus, xPlus = pad_model(v, pad, pad, free_surface=free_surface)
    mp, _, _ = pad_model(refl, pad, pad, free_surface=free_surface)
    zTop = 0 if free_surface else pad
    # reflectivity lives only in the physical domain
    mask = np.zeros_like(mp)
    mask[FAT + pad:FAT + pad + nx, FAT + zTop:FAT + zTop + nz] = 1.0
    mp *= mask
    NX, NZ = vp.shape
    params = dict(nz=NZ, nx=NX, dz=d, dx=d, nts=nts, dts=dts, sub=sub, zPadMinus=zTop, zPadPlus=zPlus,
                  xPadMinus=pad, xPadPlus=xPlus, alphaCos=0.95, fMax=37.5, freeSurface=int(free_surface))
    izs = 3 - 1 + zTop + FAT                 # zSource=3 -> 20 m depth
    spacing = nx // nShot
    src = [(izs, FAT + pad + spacing // 2 + spacing * s) for s in range(nShot)]
    rec = (1, izs, 1, nx, FAT + pad, 1)      # every column at 20 m depth
    wav = ricker(nts, dts, 15.0, 0.1)[None, :]
    name = "born_lsrtm_fs_80x500x1200" if free_surface else "born_lsrtm_80x500x1200"
    return Workload(name, params, vp, wav, src, rec, mp)


def fwi_layered(nShot=32, seed=31, nts=4000, dts=0.001, sub=2):
    """400 x 1000 grid at 10 m, 50-point sponge (alpha 0.95): a seeded model of
    8 gently dipping layers (1500-4500 m/s) with a salt-like 4500 m/s blob
    (the true model), and that model smoothed by a 500 m Gaussian (the
    starting model, returned as `vel`; the true one as `model`).  nShot
    sources and 1000 receivers at 20 m depth, nts=4000 at 1 ms (sub 2,
    Courant 0.225), 15 Hz Ricker."""
    rng = np.random.default_rng(seed)
    nz, nx, d, pad = 400, 1000, 10.0, 50
    zz, xx = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(nx, dtype=np.float64))
    tops = np.sort(rng.uniform(0.08, 0.92, 7)) * nz
    slopes = rng.uniform(-0.08, 0.08, 7)
    v = np.full((nx, nz), 1500.0)
    for k in range(7):
        below = zz >= tops[k] + slopes[k] * (xx - nx / 2)
        v[below] = 1500.0 + 3000.0 * (k + 1) / 7.0
    cz, cx = rng.uniform(0.35, 0.6) * nz, rng.uniform(0.3, 0.7) * nx
    blob = ((zz - cz) / (0.12 * nz)) ** 2 + ((xx - cx) / (0.15 * nx)) ** 2 <= 1.0
    v[blob] = 4500.0
    v0 = _smooth(v, 50.0)                    # 500 m Gaussian
    vt, zPlus, xPlus = pad_model(v.astype(np.float32), pad, pad)
    vs, _, _ = pad_model(v0.astype(np.float32), pad, pad)
    NX, NZ = vs.shape
    params = dict(nz=NZ, nx=NX, dz=d, dx=d, nts=nts, dts=dts, sub=sub, zPadMinus=pad, zPadPlus=zPlus,
                  xPadMinus=pad, xPadPlus=xPlus, alphaCos=0.95, fMax=37.5)
    izs = 3 - 1 + pad + FAT                  # 20 m depth
    spacing = nx // nShot
    src = [(izs, FAT + pad + spacing // 2 + spacing * s) for s in range(nShot)]
    rec = (1, izs, 1, nx, FAT + pad, 1)
    wav = ricker(nts, dts, 15.0, 0.1)[None, :]
    return Workload(f"fwi_layered_{nShot}x400x1000", params, vs, wav, src, rec, vt)


def serving_models(seed=50):
    """Four resident velocity models for the serving mix, grids from 200 x 400
    to 1000 x 3000 at 10 m with nts from 4000 down to 1000: each a
    1500-4000 m/s gradient with +/-5 % smoothed random perturbations, a
    50-point sponge, receivers every column at 20 m depth, 15 Hz Ricker.
    Returns [(id, params, vel, receivers, wavelet, source x range)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k, (nz, nx, nts, dts, sub) in enumerate([(200, 400, 4000, 0.002, 2), (500, 1200, 2000, 0.002, 2),
                                                 (800, 2000, 2000, 0.002, 2), (1000, 3000, 1000, 0.002, 2)]):
        d, pad = 10.0, 50
        z = np.arange(nz) / (nz - 1)
        v = np.broadcast_to(1500.0 + 2500.0 * z, (nx, nz)).astype(np.float64)
        pert = _smooth(rng.standard_normal((nx, nz)), 20.0)
        pert /= np.abs(pert).max()
        v = (v * (1.0 + 0.05 * pert)).astype(np.float32)
        vp, zPlus, xPlus = pad_model(v, pad, pad)
        NX, NZ = vp.shape
        params = dict(nz=NZ, nx=NX, dz=d, dx=d, nts=nts, dts=dts, sub=sub, zPadMinus=pad, zPadPlus=zPlus,
                      xPadMinus=pad, xPadPlus=xPlus, alphaCos=0.95, fMax=37.5)
        izs = 3 - 1 + pad + FAT
        rec = (1, izs, 1, nx, FAT + pad, 1)
        out.append((f"m{k}_{nz}x{nx}", params, vp, rec, ricker(nts, dts, 15.0, 0.1), (izs, FAT + pad, FAT + pad + nx)))
    return out


def nonlinear_small(seed=7, nts=1000, dts=0.004, sub=4):
    """200 x 400 grid at 10 m, 50-point sponge (alpha 0.95), 1500-3500 m/s
    gradient plus one Gaussian anomaly (+/-300 m/s, sigma 200 m), one source
    at 20 m depth in the centre, 400 receivers at 20 m every 10 m, nts=1000
    at 4 ms (sub 4)."""
    rng = np.random.default_rng(seed)
    nz, nx, d, pad = 200, 400, 10.0, 50
    zz, xx = np.meshgrid(np.arange(nz) * d, np.arange(nx) * d)
    v = 1500.0 + 2000.0 * zz / ((nz - 1) * d)
    cz, cx = rng.uniform(0.3, 0.7) * nz * d, rng.uniform(0.3, 0.7) * nx * d
    amp = 300.0 * rng.choice([-1.0, 1.0])
    v = (v + amp * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 200.0 ** 2))).astype(np.float32)
    vp, zPlus, xPlus = pad_model(v, pad, pad)
    NX, NZ = vp.shape
    params = dict(nz=NZ, nx=NX, dz=d, dx=d, nts=nts, dts=dts, sub=sub, zPadMinus=pad, zPadPlus=zPlus,
                  xPadMinus=pad, xPadPlus=xPlus, alphaCos=0.95, fMax=37.5)
    izs = 3 - 1 + pad + FAT
    src = [(izs, FAT + pad + nx // 2)]
    rec = (1, izs, 1, nx, FAT + pad, 1)
    wav = ricker(nts, dts, 15.0, 0.1)[None, :]
    return Workload("nonlinear_1x200x400", params, vp, wav, src, rec, np.zeros_like(vp))


@dataclass
class WindowedWorkload:
    """Shots on model windows of uneven size (the batched form): window k is
    columns [x0[k], x0[k] + nx[k]) of a wide model, padded on its own."""
    name: str
    params: dict          # keys every window shares (nz, nx: those of window 0)
    vels: list            # padded [NX_k][NZ] per window
    src: list             # per window (iz, ix) on its padded grid
    rec: list             # per window (nzDev, ozDev, dzDev, nxDev, oxDev, dxDev)
    x0: list              # first model column of each window
    nx: list              # physical width of each window
    wavelet: np.ndarray   # [1][nts]

    @property
    def nShot(self):
        return len(self.vels)

    @property
    def fine_steps(self):
        return (self.params["nts"] - 1) * self.params["sub"]

    def interior(self, k):
        NX, NZ = self.vels[k].shape
        return (NZ - 2 * FAT) * (NX - 2 * FAT)

    def window_params(self, k):
        """The single-window parameters of window k (a plain context)."""
        NX, NZ = self.vels[k].shape
        return dict(self.params, nz=NZ, nx=NX, xPadPlus=NX - 2 * FAT - self.params["xPadMinus"] - self.nx[k])


def batched_windows(nShot=128, seed=303, nts=3000, dts=0.002, sub=2, nx_range=(300, 1500), nz=400):
    """Shots on windows of uneven aperture: nz = 400 at 10 m, window widths
    drawn uniformly from nx_range, each window centred on its source, sources
    every 40 columns along a wide model.  Velocity: 1500 m/s at the top,
    growing with depth to a bottom value that varies smoothly and randomly
    along x (2800-3800 m/s), plus 24 seeded Gaussian lenses (+/-15 %,
    radius 100-300 m), clipped to 1400-4400 m/s.  50-point sponge (alpha
    0.95), source and receivers (every column of the window) at 20 m depth,
    15 Hz Ricker, nts at dts with sub 2 (Courant <= 0.44)."""
    rng = np.random.default_rng(seed)
    d, pad, step, margin = 10.0, 50, 40, 800
    widths = rng.integers(nx_range[0], nx_range[1] + 1, nShot)
    xs = margin + step * np.arange(nShot)
    nxFull = int(xs[-1] + margin)
    bottom = _smooth(rng.standard_normal(nxFull), 200.0)
    bottom = 3300.0 + 500.0 * bottom / np.abs(bottom).max()
    z = np.arange(nz) / (nz - 1)
    v = 1500.0 + (bottom[:, None] - 1500.0) * z[None, :]
    zz, xx = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(nxFull, dtype=np.float64))
    for _ in range(24):
        cz, cx = rng.uniform(0.1, 0.9) * nz, rng.uniform(0, nxFull)
        r = rng.uniform(10.0, 30.0)
        amp = rng.uniform(-0.15, 0.15)
        v *= 1.0 + amp * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * r * r))
    v = np.clip(v, 1400.0, 4400.0).astype(np.float32)
    vels, src, rec, x0s = [], [], [], []
    zPlus = pad_plus(nz, pad)
    izs = 3 - 1 + pad + FAT
    for k in range(nShot):
        w = int(widths[k])
        x0 = int(xs[k] - w // 2)
        vp, _, _ = pad_model(v[x0:x0 + w], pad, pad)
        vels.append(vp)
        x0s.append(x0)
        src.append((izs, FAT + pad + (int(xs[k]) - x0)))
        rec.append((1, izs, 1, w, FAT + pad, 1))
    NX0, NZ = vels[0].shape
    params = dict(nz=NZ, nx=NX0, dz=d, dx=d, nts=nts, dts=dts, sub=sub, zPadMinus=pad, zPadPlus=zPlus,
                  xPadMinus=pad, xPadPlus=NX0 - 2 * FAT - pad - int(widths[0]), alphaCos=0.95, fMax=37.5)
    wav = ricker(nts, dts, 15.0, 0.1)[None, :]
    return WindowedWorkload(f"windows_{nShot}x{nz}x{nx_range[0]}-{nx_range[1]}", params, vels, src, rec, x0s,
                            [int(w) for w in widths], wav)

"""Writes the geometry fixtures of the two benchmark configurations.

The reference's tests hold no golden vectors, so these are computed here from
the reference's integer and float rules, restated with numpy scalar arithmetic
(independent of the oracle's C code and of the engine's host code):

* padding (F/src/padFileGpuMain.cpp:49-66): padPlus = ceil((2 pad + n)/16)*16 - n - pad;
* grid constructor (F/src/deviceGpu.cpp:142-199): device (iz, ix) -> corners
  b, b+1, b+nz, b+nz+1 with b = ix*nz + iz and weights 1, 0, 0, 0;
* coordinate constructor (F/src/deviceGpu.cpp:8-84): w = float((x - o)/d),
  truncation, w <- float(1.0 - double(w - trunc)), corner weights formed in
  double where the C++ promotes to double (`1.0 - w`), float otherwise;
  a dipole adds 4 corners with negated weights at (z + zShift, x + xShift);
* unique points in first-seen order over (device, corner)
  (F/src/deviceGpu.cpp:201-224).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FAT = 5


def pad_plus(n, pad):
    return -(-(2 * pad + n) // 16) * 16 - n - pad


def grid_ctor(nzDev, ozDev, dzDev, nxDev, oxDev, dxDev, nz):
    idx, w = [], []
    for ix in range(nxDev):
        for iz in range(nzDev):
            b = (oxDev + ix * dxDev) * nz + (ozDev + iz * dzDev)
            idx += [b, b + 1, b + nz, b + nz + 1]
            w += [1.0, 0.0, 0.0, 0.0]
    return np.array(idx, np.int32), np.array(w, np.float32)


def _corners(z, x, oz, dz, ox, dx, nz, sign):
    f32, f64 = np.float32, np.float64
    wz = f32(f32(f32(z) - f32(oz)) / f32(dz))
    wx = f32(f32(f32(x) - f32(ox)) / f32(dx))
    zr, xr = int(wz), int(wx)
    wz = f32(1.0 - f64(f32(wz - f32(zr))))
    wx = f32(1.0 - f64(f32(wx - f32(xr))))
    b = xr * nz + zr
    if sign > 0:
        w = [f32(wz * wx), f32((1.0 - f64(wz)) * f64(wx)), f32(f64(wz) * (1.0 - f64(wx))),
             f32((1.0 - f64(wz)) * (1.0 - f64(wx)))]
    else:
        w = [f32(-1.0 * f64(wz) * f64(wx)), f32(-1.0 * (1.0 - f64(wz)) * f64(wx)),
             f32(-1.0 * f64(wz) * (1.0 - f64(wx))), f32(-1.0 * (1.0 - f64(wz)) * (1.0 - f64(wx)))]
    return [b, b + 1, b + nz, b + nz + 1], w


def coords_ctor(zs, xs, oz, dz, ox, dx, nz, dipole=0, zShift=0.0, xShift=0.0):
    idx, w = [], []
    for z, x in zip(zs, xs):
        i, ww = _corners(z, x, oz, dz, ox, dx, nz, 1)
        idx += i; w += ww
        if dipole == 1:
            i, ww = _corners(np.float32(np.float32(z) + np.float32(zShift)),
                             np.float32(np.float32(x) + np.float32(xShift)), oz, dz, ox, dx, nz, -1)
            idx += i; w += ww
    return np.array(idx, np.int32), np.array(w, np.float32)


def unique_first_seen(idx):
    seen, out = set(), []
    for v in idx.tolist():
        if v not in seen:
            seen.add(v)
            out.append(v)
    return np.array(out, np.int32)


def config(name, nz, nx, pad, src_iz, src_ix, rec):
    NZ = nz + pad + pad_plus(nz, pad) + 2 * FAT
    NX = nx + pad + pad_plus(nx, pad) + 2 * FAT
    out = {"NZ": np.int32(NZ), "NX": np.int32(NX)}
    ri, rw = grid_ctor(*rec, NZ)
    out["rec_gridIdx"], out["rec_weight"], out["rec_regPos"] = ri, rw, unique_first_seen(ri)
    out["src_regPos"] = np.array([unique_first_seen(grid_ctor(1, src_iz, 1, 1, ix, 1, NZ)[0]) for ix in src_ix],
                                 np.int32)
    # irregular receivers on the same line, 0.3 cell deeper and staggered by a
    # quarter cell, physical origin 0 and spacing 10 m
    nrec = rec[3]
    zs = np.full(nrec, (rec[1] + 0.3) * 10.0, np.float32)
    xs = ((rec[4] + np.arange(nrec) * rec[5]) * 10.0 + 2.5).astype(np.float32)
    ci, cw = coords_ctor(zs, xs, 0.0, 10.0, 0.0, 10.0, NZ)
    out.update(coord_z=zs, coord_x=xs, coord_gridIdx=ci, coord_weight=cw, coord_regPos=unique_first_seen(ci))
    di, dw = coords_ctor(zs[:50], xs[:50], 0.0, 10.0, 0.0, 10.0, NZ, dipole=1, zShift=20.0, xShift=0.0)
    out.update(dip_gridIdx=di, dip_weight=dw, dip_regPos=unique_first_seen(di))
    np.savez_compressed(os.path.join(HERE, f"geometry_{name}.npz"), **out)
    print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    # nonlinear, 200 x 400 at 10 m, pad 50: one source at 20 m depth in the
    # centre, 400 receivers at 20 m depth every 10 m
    config("nonlinear_1x200x400", 200, 400, 50, 57, [255], (1, 57, 1, 400, 55, 1))
    # Born benchmark, 500 x 1200 at 10 m, pad 50: 80 sources every 15 cells,
    # 1200 receivers on every column at 20 m depth
    config("born_lsrtm_80x500x1200", 500, 1200, 50, 57, [FAT + 50 + 7 + 15 * s for s in range(80)],
           (1, 57, 1, 1200, 55, 1))

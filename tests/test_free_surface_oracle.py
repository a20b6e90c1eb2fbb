"""Pinning the oracle's free surface (CPU only).

With freeSurface = 1 the reference zeroes the surface row (iz = FAT) of the
current slice and mirrors the rows above it with the opposite sign before
every forward step (setFsConditionFwdGpu, F/src/kernelsGpu.cu:555-567), drops
the top sponge (dampCosineEdgeFs, :221-239) and uses a modified adjoint
stepper (stepAdjFsGpu, :613-700).

* Forward, by the method of images: on a grid mirrored about the surface
  row, with the velocity mirrored and an opposite-signed image source, the
  field is odd about that row at every step, so the free-surface field
  equals it on the lower half exactly (this holds with the sponge and the
  grid's own boundaries, which are mirrored too).
* Adjoint: brute-force transposition of the nonlinear and Born operators
  with the free surface, and FP64 dot tests.
"""
import numpy as np
import pytest

from cases import rel_l2
from oracle import oracle as O

FAT = 5


def _fs_grid(nz=30, nx=34, nts=40, sub=2, minPad=6, alpha=0.9, nExt=1):
    return O.Grid(nz=nz, nx=nx, nts=nts, sub=sub, minPad=minPad, dz=10.0, dx=10.0, dts=0.002, alphaCos=alpha,
                  nExt=nExt, freeSurface=1)


@pytest.mark.parametrize("prec", ["d", "f"])
def test_free_surface_equals_the_method_of_images(prec):
    g = _fs_grid()
    rng = np.random.default_rng(1)
    vel = 2000.0 + 300.0 * rng.random((g.nx, g.nz))
    I = g.nz - 2 * FAT                              # interior rows, the surface row first
    nzb = 2 * I - 1 + 2 * FAT
    c = FAT + I - 1                                 # the surface row on the mirrored grid
    big = O.Grid(nz=nzb, nx=g.nx, nts=g.nts, sub=g.sub, minPad=g.minPad, dz=g.dz, dx=g.dx, dts=g.dts,
                 alphaCos=g.alphaCos, freeSurface=0)
    vb = np.zeros((g.nx, nzb))
    for k in range(I):
        vb[:, c + k] = vel[:, FAT + k]
        vb[:, c - k] = vel[:, FAT + k]
    vb[:, :FAT] = vb[:, FAT:FAT + 1]
    vb[:, -FAT:] = vb[:, -FAT - 1:-FAT]
    zs, xs = 9, 14
    w = np.sin(np.arange(g.nts) * 0.35) * np.exp(-0.06 * np.arange(g.nts))
    src = [O.Device.ints([zs], [xs], g.nz, g.nx, g.nts)]
    recz = np.array([FAT, FAT + 1, FAT + 3, 12, 20, g.nz - FAT - 1])
    recx = np.array([7, 11, 16, 20, 25, 28])
    rec = [O.Device.ints(recz, recx, g.nz, g.nx, g.nts)]
    d_fs = O.nl_forward(g, vel, src, rec, w[None, None, :], prec)
    srcb = [O.Device.ints([c + zs - FAT, c - (zs - FAT)], [xs, xs], nzb, g.nx, g.nts)]
    recb = [O.Device.ints(c + recz - FAT, recx, nzb, g.nx, g.nts)]
    d_big = O.nl_forward(big, vb, srcb, recb, np.stack([w, -w])[None], prec)
    print(f"{prec}: |d| {np.abs(d_fs).max():.3e}, relL2 free surface vs images {rel_l2(d_fs, d_big):.2e}")
    assert np.abs(d_fs).max() > 0
    assert np.all(d_fs[0, 0] == 0)                  # the surface row itself
    assert rel_l2(d_fs, d_big) < (1e-13 if prec == "d" else 1e-6)
    # and without the free surface the same grid gives a different answer
    g0 = _fs_grid()
    g0.freeSurface = 0
    assert rel_l2(O.nl_forward(g0, vel, src, rec, w[None, None, :], prec), d_fs) > 0.1


def _nl_matrices(g, vel, src, rec):
    nm = g.nts
    F = np.stack([O.nl_forward(g, vel, src, rec, np.eye(nm)[i][None, None], "d").ravel() for i in range(nm)], 1)
    nd = F.shape[0]
    A = np.stack([O.nl_adjoint(g, vel, src, rec, np.eye(nd)[k].reshape(len(src), -1, g.nts), 1, "d").ravel()
                  for k in range(nd)], 1)
    return F, A


def test_free_surface_nonlinear_transpose():
    g = _fs_grid(nz=22, nx=22, nts=6, minPad=3)
    rng = np.random.default_rng(2)
    vel = 2000.0 + rng.uniform(-200, 200, (g.nx, g.nz))
    src = [O.Device.grid(1, 7, 1, 1, 9, 1, g.nz, g.nx, g.nts), O.Device.grid(1, 6, 1, 1, 13, 1, g.nz, g.nx, g.nts)]
    rec = [O.Device.ints([FAT + 1, FAT + 2, 9, 12], [10, 12, 9, 14], g.nz, g.nx, g.nts)]
    F, A = _nl_matrices(g, vel, src, rec)
    assert np.abs(F).max() > 0
    np.testing.assert_allclose(A, F.T, atol=1e-12 * np.abs(F).max())


def test_receivers_on_the_surface_row_break_transposition():
    """A reference quirk kept as is: the forward field is zero on the surface
    row, so a receiver there records zeros, but stepAdjFsGpu leaves that row
    to the injected adjoint traces, which then reach the rows below through
    the stencil.  The pair stays adjoint for receivers below the surface."""
    g = _fs_grid(nz=22, nx=22, nts=6, minPad=3)
    rng = np.random.default_rng(2)
    vel = 2000.0 + rng.uniform(-200, 200, (g.nx, g.nz))
    src = [O.Device.grid(1, 7, 1, 1, 9, 1, g.nz, g.nx, g.nts)]
    rec = [O.Device.ints([FAT, 9], [8, 9], g.nz, g.nx, g.nts)]
    F, A = _nl_matrices(g, vel, src, rec)
    surface = np.arange(F.shape[0]) < g.nts                 # rows of the surface receiver's trace
    assert np.all(F[surface] == 0) and np.abs(A[:, surface]).max() > 0
    np.testing.assert_allclose(A[:, ~surface], F[~surface].T, atol=1e-12 * np.abs(F).max())


def test_free_surface_born_transpose_and_dot_products():
    g = _fs_grid(nz=22, nx=22, nts=6, minPad=3)
    rng = np.random.default_rng(3)
    vel = 2000.0 + rng.uniform(-200, 200, (g.nx, g.nz))
    src = [O.Device.grid(1, 7, 1, 1, 9, 1, g.nz, g.nx, g.nts)]
    rec = [O.Device.grid(1, 6, 1, 5, 7, 2, g.nz, g.nx, g.nts)]   # all inside the FAT ring (ix <= 15)
    wav = [rng.standard_normal((1, g.nts))]
    nm = g.nx * g.nz
    F = np.stack([O.born_forward(g, vel, src, wav, rec, np.eye(nm)[i].reshape(g.nx, g.nz), "d").ravel()
                  for i in range(nm)], 1)
    nd = F.shape[0]
    A = np.stack([O.born_adjoint(g, vel, src, wav, rec, np.eye(nd)[k].reshape(1, -1, g.nts), "d").ravel()
                  for k in range(nd)], 1)
    assert np.abs(F).max() > 0
    np.testing.assert_allclose(A, F.T, atol=1e-12 * np.abs(F).max())
    g5 = _fs_grid(nExt=3)
    vel5 = 2000.0 + rng.uniform(-200, 200, (g5.nx, g5.nz))
    src5 = [O.Device.grid(1, 8, 1, 1, 12, 1, g5.nz, g5.nx, g5.nts)]
    rec5 = [O.Device.grid(1, FAT + 1, 1, 10, 8, 2, g5.nz, g5.nx, g5.nts)]
    w5 = [rng.standard_normal((1, g5.nts))]
    d = rng.standard_normal((1, rec5[0].nDev, g5.nts))
    for ext in (False, True):
        m = rng.standard_normal((3, g5.nx, g5.nz) if ext else (g5.nx, g5.nz))
        a = np.vdot(O.born_forward(g5, vel5, src5, w5, rec5, m, "d", ext).ravel(), d.ravel())
        b = np.vdot(m.ravel(), O.born_adjoint(g5, vel5, src5, w5, rec5, d, "d", ext).ravel())
        assert abs(a - b) <= 1e-12 * abs(b)

"""Pinning the oracle's offset-extended WEMVA operator (CPU only).

wemvaExtShotsGpu (F/src/wemvaExtGpu.cpp, F/src/wemvaExtShotsGpuFunctions.cu)
maps a velocity perturbation dv to the perturbation of the extended image
of the observed data: leg 1 perturbs the source wavefield (imaged against
the receiver wavefield), leg 2 the receiver wavefield (imaged against the
source wavefield's second time derivative).  With sub = 1 it is the exact
velocity derivative of the extended Born adjoint image of the data at image
points away from dv (where the image's own 2/v^3 and v^2 dtw^2 scalings do
not move); each leg is fitted separately.  The adjoint is pinned by
brute-force transposition and FP64 dot tests.
"""
import numpy as np
import pytest

from cases import make_case, rel_l2, ricker
from oracle import oracle as O

FAT = 5


def _fit(nts, dts, sub, seed=91):
    c = make_case(seed=seed, nShot=1, nts=nts, dts=dts, sub=sub, nExt=5)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    rng = np.random.default_rng(92)
    w = ricker(c.nts, c.dts, 15.0, 0.08).astype(np.float64)
    dobs = rng.standard_normal((1, rec[0].nDev, c.nts)) * np.hanning(c.nts)[None, None]
    dv = np.zeros((c.nx, c.nz))
    dv[FAT + 10:c.nx - FAT - 10, 26:36] = 20.0 * rng.standard_normal((c.nx - 2 * FAT - 20, 10))
    eps = 1e-3
    fd = (O.born_adjoint(g, c.vel + eps * dv, src, [w[None]], rec, dobs, "d", ext=True)
          - O.born_adjoint(g, c.vel - eps * dv, src, [w[None]], rec, dobs, "d", ext=True)) / (2 * eps)
    t1 = O.wemva_forward(g, c.vel, src, [w[None]], rec, dobs, dv, 1, 0, prec="d")
    t2 = O.wemva_forward(g, c.vel, src, [w[None]], rec, dobs, dv, 0, 1, prec="d")
    keep = np.ones((c.nx, c.nz), bool)            # image points whose scalings do not move
    keep[FAT + 10 - 3:c.nx - FAT - 10 + 3, 25:37] = False
    sel = np.broadcast_to(keep, fd.shape)
    A = np.stack([t1[sel], t2[sel]], 1)
    coef, *_ = np.linalg.lstsq(A, fd[sel], rcond=None)
    resid = np.linalg.norm(A @ coef - fd[sel]) / np.linalg.norm(fd[sel])
    print(f"sub {sub}: finite difference ~ {coef[0]:.6f} leg1 + {coef[1]:.6f} leg2, residual {resid:.2e}")
    assert np.abs(t1).max() > 0 and np.abs(t2).max() > 0
    return coef, resid


def test_each_leg_is_the_velocity_derivative_of_the_extended_image():
    coef, resid = _fit(240, 0.001, 1)
    np.testing.assert_allclose(coef, [1.0, 1.0], atol=1e-4)
    assert resid < 1e-4


def test_with_two_fine_steps_per_sample_the_derivative_is_close():
    """sub = 2: 0.7 % off on this case (the coarse second derivative of the
    source wavefield); the stated 3 % fails a dropped leg or a sign."""
    coef, resid = _fit(120, 0.002, 2)
    np.testing.assert_allclose(coef, [1.0, 1.0], atol=0.03)
    assert resid < 0.03


@pytest.mark.parametrize("legs", [(1, 0), (0, 1)])
def test_brute_force_transpose(legs):
    g = O.Grid(nz=24, nx=24, nts=5, sub=2, minPad=3, dz=10.0, dx=10.0, dts=0.002, alphaCos=0.9, nExt=3)
    rng = np.random.default_rng(3)
    vel = np.full((g.nx, g.nz), 2000.0) + rng.uniform(-200, 200, (g.nx, g.nz))
    src = [O.Device.grid(1, 9, 1, 1, 9, 1, g.nz, g.nx, g.nts)]
    rec = [O.Device.grid(1, 8, 1, 6, 7, 2, g.nz, g.nx, g.nts)]
    wav = [rng.standard_normal((1, g.nts))]
    dobs = rng.standard_normal((1, rec[0].nDev, g.nts))
    nm = g.nx * g.nz
    F = np.stack([O.wemva_forward(g, vel, src, wav, rec, dobs, np.eye(nm)[i].reshape(g.nx, g.nz), *legs,
                                  prec="d").ravel() for i in range(nm)], 1)
    ni = F.shape[0]
    A = np.stack([O.wemva_adjoint(g, vel, src, wav, rec, dobs, np.eye(ni)[k].reshape(g.nExt, g.nx, g.nz), *legs,
                                  prec="d").ravel() for k in range(ni)], 1)
    assert np.abs(F).max() > 0
    np.testing.assert_allclose(A, F.T, atol=1e-12 * np.abs(F).max())


def test_dot_products_fp64_and_float_tracks_fp64():
    c = make_case(seed=93, nShot=2, nExt=5, sub=2, dts=0.002)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    rng = np.random.default_rng(94)
    wav = [c.wavelet[None].astype(np.float64)]
    dobs = rng.standard_normal((c.nShot, rec[0].nDev, c.nts))
    m = rng.standard_normal((c.nx, c.nz))
    im = rng.standard_normal((5, c.nx, c.nz))
    for legs in [(1, 0), (0, 1), (1, 1)]:
        a = np.vdot(O.wemva_forward(g, c.vel, src, wav, rec, dobs, m, *legs, prec="d").ravel(), im.ravel())
        b = np.vdot(m.ravel(), O.wemva_adjoint(g, c.vel, src, wav, rec, dobs, im, *legs, prec="d").ravel())
        print(f"legs {legs}: <Fm,i> {a:.12e} <m,F'i> {b:.12e}")
        assert abs(a - b) <= 1e-12 * abs(b)
    f32 = np.float32
    fF = O.wemva_forward(g, c.vel, src, [wav[0].astype(f32)], rec, dobs.astype(f32), m.astype(f32))
    fD = O.wemva_forward(g, c.vel, src, wav, rec, dobs, m, prec="d")
    aF = O.wemva_adjoint(g, c.vel, src, [wav[0].astype(f32)], rec, dobs.astype(f32), im.astype(f32))
    aD = O.wemva_adjoint(g, c.vel, src, wav, rec, dobs, im, prec="d")
    print(f"float vs FP64: forward {rel_l2(fF, fD):.2e}, adjoint {rel_l2(aF, aD):.2e}")
    assert rel_l2(fF, fD) < 1e-4 and rel_l2(aF, aD) < 1e-4

"""Pinning the oracle's offset-extended tomographic operator (CPU only).

tomoExtShotsGpu (F/src/tomoExtGpu.cpp, F/src/tomoExtShotsGpuFunctions.cu) maps
a velocity perturbation dv to data about a fixed extended reflectivity r:
leg 1 scatters the source wavefield's second time derivative by r and then
by dv, leg 2 by dv and then by r.  With one fine step per coarse sample
(sub = 1) the coarse second derivative is the fine one, and the operator is
the exact velocity derivative of the extended Born data for perturbations
away from the sources and from r's support; each leg is fitted separately,
so a wrong sign, scale or offset index in either leg fails.  The adjoint is
pinned by brute-force transposition and FP64 dot tests, the coarse second
derivative by quadratics.
"""
import numpy as np
import pytest

from cases import make_case, rel_l2, ricker
from oracle import oracle as O

FAT = 5


def _setup(nts, dts, sub, nExt=5, seed=71):
    c = make_case(seed=seed, nShot=1, nts=nts, dts=dts, sub=sub, nExt=nExt)
    rng = np.random.default_rng(72)
    w = ricker(c.nts, c.dts, 15.0, 0.08).astype(np.float64)
    r = np.zeros((nExt, c.nx, c.nz))
    r[:, FAT + 8:c.nx - FAT - 8, 45:48] = rng.standard_normal((nExt, c.nx - 2 * FAT - 16, 3))
    dv = np.zeros((c.nx, c.nz))
    dv[FAT + 10:c.nx - FAT - 10, 26:36] = 20.0 * rng.standard_normal((c.nx - 2 * FAT - 20, 10))
    return c, w, r, dv


def _fd_and_legs(c, w, r, dv, eps=1e-3):
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    fd = (O.born_forward(g, c.vel + eps * dv, src, [w[None]], rec, r, "d", ext=True)
          - O.born_forward(g, c.vel - eps * dv, src, [w[None]], rec, r, "d", ext=True)) / (2 * eps)
    t1 = O.tomo_forward(g, c.vel, src, [w[None]], rec, dv, r, leg1=1, leg2=0, prec="d")
    t2 = O.tomo_forward(g, c.vel, src, [w[None]], rec, dv, r, leg1=0, leg2=1, prec="d")
    return fd, t1, t2


def test_coarse_second_derivative_of_quadratics():
    g = O.Grid(nz=16, nx=14, nts=9, sub=2, minPad=3, dz=10.0, dx=10.0, dts=0.004, alphaCos=0.9, nExt=1)
    t = np.arange(g.nts) * g.dts
    rng = np.random.default_rng(1)
    a, b, q = rng.standard_normal(3)
    W = np.zeros((g.nts, g.nx, g.nz))
    W[:, FAT:-FAT, FAT:-FAT] = (a + b * t + q * t * t)[:, None, None]
    D = O.dt2_coarse(g, W, "d")
    inner = D[1:-1, FAT:-FAT, FAT:-FAT]
    np.testing.assert_allclose(inner, 2 * q, rtol=1e-6, atol=1e-6 * abs(q))
    d2 = g.dts ** 2
    w0, w1, wl, wm = W[0, 7, 7], W[1, 7, 7], W[-1, 7, 7], W[-2, 7, 7]
    assert D[0, 7, 7] == pytest.approx((w1 - 2 * w0) / d2, rel=1e-12)       # zero before t = 0
    assert D[-1, 7, 7] == pytest.approx((wm - 2 * wl) / d2, rel=1e-12)      # zero after the last sample
    assert np.all(D[:, :FAT] == 0) and np.all(D[:, :, :FAT] == 0)


def test_each_leg_is_the_velocity_derivative_of_extended_born():
    c, w, r, dv = _setup(nts=240, dts=0.001, sub=1)
    fd, t1, t2 = _fd_and_legs(c, w, r, dv)
    A = np.stack([t1.ravel(), t2.ravel()], 1)
    coef, *_ = np.linalg.lstsq(A, fd.ravel(), rcond=None)
    resid = np.linalg.norm(A @ coef - fd.ravel()) / np.linalg.norm(fd)
    print(f"finite difference ~ {coef[0]:.6f} leg1 + {coef[1]:.6f} leg2, residual {resid:.2e}")
    assert np.abs(t1).max() > 0 and np.abs(t2).max() > 0
    np.testing.assert_allclose(coef, [1.0, 1.0], atol=1e-4)
    assert resid < 1e-4


def test_with_two_fine_steps_per_sample_the_derivative_is_approximate():
    """With sub = 2 the operator applies the second derivative to the coarse
    source wavefield (F/src/tomoExtShotsGpuFunctions.cu:1251-1360) while the
    exact derivative acts on every fine step of the time-interpolated source;
    the two differ by a sub-dependent amount that does not shrink with the
    frequency (13 % at sub = 2 on this case).  The stated tolerance 20 %
    still fails a dropped leg (55-67 %) or a sign error (190 %)."""
    c, w, r, dv = _setup(nts=120, dts=0.002, sub=2)
    fd, t1, t2 = _fd_and_legs(c, w, r, dv)
    err = rel_l2(t1 + t2, fd)
    print(f"sub 2: relL2 {err:.3f}; leg 1 alone {rel_l2(t1, fd):.3f}, leg 2 alone {rel_l2(t2, fd):.3f}")
    assert err < 0.2


def _tiny(nExt=3):
    return O.Grid(nz=26, nx=26, nts=6, sub=2, minPad=4, dz=10.0, dx=10.0, dts=0.002, alphaCos=0.9, nExt=nExt)


@pytest.mark.parametrize("legs", [(1, 0), (0, 1)])
def test_brute_force_transpose(legs):
    g = _tiny()
    rng = np.random.default_rng(2)
    vel = np.full((g.nx, g.nz), 2000.0) + rng.uniform(-200, 200, (g.nx, g.nz))
    src = [O.Device.grid(1, 9, 1, 1, 9, 1, g.nz, g.nx, g.nts)]
    rec = [O.Device.grid(1, 8, 1, 7, 7, 2, g.nz, g.nx, g.nts)]
    wav = [rng.standard_normal((1, g.nts))]
    r = rng.standard_normal((g.nExt, g.nx, g.nz))
    nm = g.nx * g.nz
    F = np.stack([O.tomo_forward(g, vel, src, wav, rec, np.eye(nm)[i].reshape(g.nx, g.nz), r, *legs, prec="d").ravel()
                  for i in range(nm)], 1)
    nd = F.shape[0]
    A = np.stack([O.tomo_adjoint(g, vel, src, wav, rec, np.eye(nd)[k].reshape(1, -1, g.nts), r, *legs,
                                 prec="d").ravel() for k in range(nd)], 1)
    assert np.abs(F).max() > 0
    np.testing.assert_allclose(A, F.T, atol=1e-12 * np.abs(F).max())


def test_dot_products_fp64_and_float_tracks_fp64():
    c = make_case(seed=73, nShot=2, nExt=5, sub=2, dts=0.002)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    rng = np.random.default_rng(74)
    wav = [c.wavelet[None].astype(np.float64)]
    r = rng.standard_normal((5, c.nx, c.nz))
    m = rng.standard_normal((c.nx, c.nz))
    d = rng.standard_normal((c.nShot, rec[0].nDev, c.nts))
    for legs in [(1, 0), (0, 1), (1, 1)]:
        a = np.vdot(O.tomo_forward(g, c.vel, src, wav, rec, m, r, *legs, prec="d").ravel(), d.ravel())
        b = np.vdot(m.ravel(), O.tomo_adjoint(g, c.vel, src, wav, rec, d, r, *legs, prec="d").ravel())
        print(f"legs {legs}: <Fm,d> {a:.12e} <m,F'd> {b:.12e}")
        assert abs(a - b) <= 1e-12 * abs(b)
    fF = O.tomo_forward(g, c.vel, src, [wav[0].astype(np.float32)], rec, m.astype(np.float32), r.astype(np.float32))
    fD = O.tomo_forward(g, c.vel, src, wav, rec, m, r, prec="d")
    aF = O.tomo_adjoint(g, c.vel, src, [wav[0].astype(np.float32)], rec, d.astype(np.float32), r.astype(np.float32))
    aD = O.tomo_adjoint(g, c.vel, src, wav, rec, d, r, prec="d")
    print(f"float vs FP64: forward {rel_l2(fF, fD):.2e}, adjoint {rel_l2(aF, aD):.2e}")
    assert rel_l2(fF, fD) < 1e-4 and rel_l2(aF, aD) < 1e-4

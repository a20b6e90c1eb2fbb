"""Pinning the oracle's time-lag tomographic and WEMVA operators
(extension=time).  CPU only.

As for subsurface offsets, with one fine step per coarse sample (sub = 1)
each operator is the exact velocity derivative of a time-lag extended Born
quantity, for a perturbation dv away from the sources and from where the
extended model's own scaling (2/v^3 v^2 dtw^2) sits:

* tomographic: the time-lag Born data about a fixed extended reflectivity r;
* WEMVA: the time-lag Born adjoint image of the observed data.

Each leg is fitted separately, so a wrong lag direction, bound or scale in
either leg fails.  The adjoints are pinned by brute-force transposition on a
tiny grid whose lag bounds cut at both ends of the time axis and by FP64 dot
tests.  The reference's WEMVA has no free-surface routines: a free surface
leaves it unchanged.
"""
import numpy as np
import pytest

from cases import make_case, rel_l2, ricker
from oracle import oracle as O

FAT = 5


def _case(nts, dts, sub, nExt=7, seed=81):
    c = make_case(seed=seed, nShot=1, nts=nts, dts=dts, sub=sub, nExt=nExt)
    rng = np.random.default_rng(82)
    w = ricker(c.nts, c.dts, 15.0, 0.08).astype(np.float64)
    dv = np.zeros((c.nx, c.nz))
    dv[FAT + 10:c.nx - FAT - 10, 26:36] = 20.0 * rng.standard_normal((c.nx - 2 * FAT - 20, 10))
    return c, w, dv, rng


def _fit(fd, t1, t2, sel=None):
    if sel is None:
        sel = np.ones(fd.shape, bool)
    A = np.stack([t1[sel], t2[sel]], 1)
    coef, *_ = np.linalg.lstsq(A, fd[sel], rcond=None)
    resid = np.linalg.norm(A @ coef - fd[sel]) / np.linalg.norm(fd[sel])
    assert np.abs(t1).max() > 0 and np.abs(t2).max() > 0
    return coef, resid


@pytest.mark.parametrize("sloth", [0, 1])
def test_tomo_legs_are_the_velocity_derivative_of_time_lag_born(sloth):
    c, w, dv, rng = _case(240, 0.001, 1)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    r = np.zeros((c.nExt, c.nx, c.nz))
    r[:, FAT + 8:c.nx - FAT - 8, 45:48] = rng.standard_normal((c.nExt, c.nx - 2 * FAT - 16, 3))
    eps = 1e-3
    fd = (O.borntime_forward(g, c.vel + eps * dv, src, [w[None]], rec, r, sloth, "d")
          - O.borntime_forward(g, c.vel - eps * dv, src, [w[None]], rec, r, sloth, "d")) / (2 * eps)
    t1 = O.tomo_forward(g, c.vel, src, [w[None]], rec, dv, r, 1, 0, "d", time=True, sloth=sloth)
    t2 = O.tomo_forward(g, c.vel, src, [w[None]], rec, dv, r, 0, 1, "d", time=True, sloth=sloth)
    coef, resid = _fit(fd, t1, t2)
    print(f"sloth {sloth}: finite difference ~ {coef[0]:.6f} leg1 + {coef[1]:.6f} leg2, residual {resid:.2e}")
    np.testing.assert_allclose(coef, [1.0, 1.0], atol=1e-4)
    assert resid < 1e-4


def test_wemva_legs_are_the_velocity_derivative_of_the_time_lag_image():
    c, w, dv, rng = _case(240, 0.001, 1, seed=83)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    dobs = rng.standard_normal((1, rec[0].nDev, c.nts)) * np.hanning(c.nts)[None, None]
    # Leg 2 moves the second time derivative from the receiver wavefield to
    # the source side; for a negative lag the sum over samples starts where
    # the source wavefield is sampled 2 |lag| samples in, which leaves a
    # boundary term in W there (3.7e-4 of the derivative on this case with
    # the Ricker's early tail).  A wavelet that is exactly zero over its
    # first samples makes W vanish there.
    w[:12] = 0.0
    eps = 1e-3
    fd = (O.borntime_adjoint(g, c.vel + eps * dv, src, [w[None]], rec, dobs, 0, "d")
          - O.borntime_adjoint(g, c.vel - eps * dv, src, [w[None]], rec, dobs, 0, "d")) / (2 * eps)
    t1 = O.wemva_forward(g, c.vel, src, [w[None]], rec, dobs, dv, 1, 0, "d", time=True)
    t2 = O.wemva_forward(g, c.vel, src, [w[None]], rec, dobs, dv, 0, 1, "d", time=True)
    keep = np.ones((c.nx, c.nz), bool)            # image points whose scalings do not move
    keep[FAT + 10 - 3:c.nx - FAT - 10 + 3, 25:37] = False
    coef, resid = _fit(fd, t1, t2, np.broadcast_to(keep, fd.shape))
    print(f"finite difference ~ {coef[0]:.6f} leg1 + {coef[1]:.6f} leg2, residual {resid:.2e}")
    np.testing.assert_allclose(coef, [1.0, 1.0], atol=1e-4)
    assert resid < 1e-4


def _tiny():
    return O.Grid(nz=24, nx=24, nts=7, sub=2, minPad=3, dz=10.0, dx=10.0, dts=0.002, alphaCos=0.9, nExt=5)


def _unit(n, i):
    e = np.zeros(n)
    e[i] = 1.0
    return e


@pytest.mark.parametrize("legs,sloth", [((1, 0), 0), ((0, 1), 1)])
def test_tomo_brute_force_transpose(legs, sloth):
    g = _tiny()
    rng = np.random.default_rng(4)
    vel = np.full((g.nx, g.nz), 2000.0) + rng.uniform(-200, 200, (g.nx, g.nz))
    src = [O.Device.grid(1, 9, 1, 1, 9, 1, g.nz, g.nx, g.nts)]
    rec = [O.Device.grid(1, 8, 1, 6, 7, 2, g.nz, g.nx, g.nts)]
    wav = [rng.standard_normal((1, g.nts))]
    r = rng.standard_normal((g.nExt, g.nx, g.nz))
    nm = g.nx * g.nz
    F = np.stack([O.tomo_forward(g, vel, src, wav, rec, _unit(nm, i).reshape(g.nx, g.nz), r, *legs, "d",
                                 time=True, sloth=sloth).ravel() for i in range(nm)], 1)
    nd = F.shape[0]
    A = np.stack([O.tomo_adjoint(g, vel, src, wav, rec, _unit(nd, k).reshape(1, -1, g.nts), r, *legs, "d",
                                 time=True, sloth=sloth).ravel() for k in range(nd)], 1)
    assert np.abs(F).max() > 0
    np.testing.assert_allclose(A, F.T, atol=1e-12 * np.abs(F).max())


@pytest.mark.parametrize("legs", [(1, 0), (0, 1)])
def test_wemva_brute_force_transpose(legs):
    g = _tiny()
    rng = np.random.default_rng(5)
    vel = np.full((g.nx, g.nz), 2000.0) + rng.uniform(-200, 200, (g.nx, g.nz))
    src = [O.Device.grid(1, 9, 1, 1, 9, 1, g.nz, g.nx, g.nts)]
    rec = [O.Device.grid(1, 8, 1, 6, 7, 2, g.nz, g.nx, g.nts)]
    wav = [rng.standard_normal((1, g.nts))]
    dobs = rng.standard_normal((1, rec[0].nDev, g.nts))
    nm = g.nx * g.nz
    F = np.stack([O.wemva_forward(g, vel, src, wav, rec, dobs, _unit(nm, i).reshape(g.nx, g.nz), *legs, "d",
                                  time=True).ravel() for i in range(nm)], 1)
    ni = F.shape[0]
    A = np.stack([O.wemva_adjoint(g, vel, src, wav, rec, dobs, _unit(ni, k).reshape(g.nExt, g.nx, g.nz), *legs,
                                  "d", time=True).ravel() for k in range(ni)], 1)
    # every lag slice is reached
    assert all(np.abs(F[j * nm:(j + 1) * nm]).max() > 0 for j in range(g.nExt))
    np.testing.assert_allclose(A, F.T, atol=1e-12 * np.abs(F).max())


def test_dot_products_fp64_and_float_tracks_fp64():
    c = make_case(seed=85, nShot=2, nExt=9, sub=3, dts=0.003, nts=50)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    rng = np.random.default_rng(86)
    wav = [c.wavelet[None].astype(np.float64)]
    r = rng.standard_normal((9, c.nx, c.nz))
    m = rng.standard_normal((c.nx, c.nz))
    d = rng.standard_normal((c.nShot, rec[0].nDev, c.nts))
    im = rng.standard_normal((9, c.nx, c.nz))
    for legs in [(1, 0), (0, 1), (1, 1)]:
        a = np.vdot(O.tomo_forward(g, c.vel, src, wav, rec, m, r, *legs, "d", time=True).ravel(), d.ravel())
        b = np.vdot(m.ravel(), O.tomo_adjoint(g, c.vel, src, wav, rec, d, r, *legs, "d", time=True).ravel())
        assert abs(a - b) <= 1e-12 * abs(b)
        a = np.vdot(O.wemva_forward(g, c.vel, src, wav, rec, d, m, *legs, "d", time=True).ravel(), im.ravel())
        b = np.vdot(m.ravel(), O.wemva_adjoint(g, c.vel, src, wav, rec, d, im, *legs, "d", time=True).ravel())
        assert abs(a - b) <= 1e-12 * abs(b)
    f32 = np.float32
    fw = [wav[0].astype(f32)]
    pairs = [
        (O.tomo_forward(g, c.vel, src, fw, rec, m.astype(f32), r.astype(f32), time=True),
         O.tomo_forward(g, c.vel, src, wav, rec, m, r, prec="d", time=True)),
        (O.tomo_adjoint(g, c.vel, src, fw, rec, d.astype(f32), r.astype(f32), time=True),
         O.tomo_adjoint(g, c.vel, src, wav, rec, d, r, prec="d", time=True)),
        (O.wemva_forward(g, c.vel, src, fw, rec, d.astype(f32), m.astype(f32), time=True),
         O.wemva_forward(g, c.vel, src, wav, rec, d, m, prec="d", time=True)),
        (O.wemva_adjoint(g, c.vel, src, fw, rec, d.astype(f32), im.astype(f32), time=True),
         O.wemva_adjoint(g, c.vel, src, wav, rec, d, im, prec="d", time=True)),
    ]
    for f, dd in pairs:
        assert rel_l2(f, dd) < 1e-4


@pytest.mark.parametrize("time", [False, True])
def test_wemva_ignores_a_free_surface(time):
    c = make_case(seed=87, nShot=1, nExt=5, sub=2, dts=0.002, nts=40)
    src, rec = c.oracle_sources(), c.oracle_receivers()
    rng = np.random.default_rng(88)
    wav = [c.wavelet[None]]
    d = rng.standard_normal((1, rec[0].nDev, c.nts)).astype(np.float32)
    m = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    g0 = c.grid
    gf = c.grid
    gf.freeSurface = 1
    np.testing.assert_array_equal(O.wemva_forward(gf, c.vel, src, wav, rec, d, m, time=time),
                                  O.wemva_forward(g0, c.vel, src, wav, rec, d, m, time=time))
    # the tomographic operator does see it
    r = rng.standard_normal((5, c.nx, c.nz)).astype(np.float32)
    assert rel_l2(O.tomo_forward(gf, c.vel, src, wav, rec, m, r, time=time),
                  O.tomo_forward(g0, c.vel, src, wav, rec, m, r, time=time)) > 1e-3

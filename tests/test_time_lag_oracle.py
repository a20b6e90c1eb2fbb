"""Pinning the oracle's time-lag extended Born operator (extension=time).  CPU only.

The reference ships no vectors for it, so the checks are what the
mathematics fixes, chosen so that a wrong lag sign, a lag step other than two
coarse samples, wrong bounds or a wrong scaling each fail one of them:

* brute-force transposition on a tiny grid whose lag bounds cut at both ends
  of the time axis (FP64), for both scalings;
* nExt == 1, and a model only at zero lag, equal the Born operator;
* a model only at lag l > 0 equals the Born operator with the wavelet delayed
  by 2 l coarse samples (its source wavefield, 2 l samples later);
* sloth=1 scales by -v^2 dtw^2 instead of 2/v^3 v^2 dtw^2; a free surface
  ignores the key, as the reference's free-surface routines do;
* dot tests on the standard case, float against FP64.
"""
import numpy as np
import pytest

from cases import make_case, rel_l2
from oracle import oracle as O

FAT = 5


def tiny(nExt, nts=6, fs=0):
    return O.Grid(nz=26, nx=26, nts=nts, sub=2, minPad=4, dz=10.0, dx=10.0, dts=0.002, alphaCos=0.9,
                  nExt=nExt, freeSurface=fs)


@pytest.mark.parametrize("nExt,nts,sloth", [(5, 6, 0), (3, 8, 1)])
def test_brute_force_transpose(nExt, nts, sloth):
    g = tiny(nExt, nts)
    rng = np.random.default_rng(3)
    vel = np.full((g.nx, g.nz), 2000.0) + rng.uniform(-200, 200, (g.nx, g.nz))
    src = [O.Device.grid(1, 9, 1, 1, 11, 1, g.nz, g.nx, g.nts)]
    rec = [O.Device.grid(1, 8, 1, 7, 7, 2, g.nz, g.nx, g.nts)]
    wav = [rng.standard_normal((1, g.nts))]
    shape = (nExt, g.nx, g.nz)
    nm = int(np.prod(shape))
    cols = []
    for i in range(nm):
        e = np.zeros(nm)
        e[i] = 1.0
        cols.append(O.borntime_forward(g, vel, src, wav, rec, e.reshape(shape), sloth, "d").ravel())
    F = np.stack(cols, 1)
    nd = F.shape[0]
    rows = []
    for k in range(nd):
        e = np.zeros(nd)
        e[k] = 1.0
        rows.append(O.borntime_adjoint(g, vel, src, wav, rec, e.reshape(1, -1, g.nts), sloth, "d").ravel())
    A = np.stack(rows, 1)
    assert np.abs(F).max() > 0
    # every lag slice reaches the data
    assert all(np.abs(F[:, j * g.nx * g.nz:(j + 1) * g.nx * g.nz]).max() > 0 for j in range(nExt))
    np.testing.assert_allclose(A, F.T, atol=1e-12 * np.abs(F).max())


def _case():
    c = make_case(seed=21, nShot=2, nts=60)
    w = c.wavelet.astype(np.float64).copy()
    w[:4] = 0.0          # zero ends, so that delaying the wavelet drops no sample
    w[-12:] = 0.0
    return c, w


@pytest.mark.parametrize("prec", ["f", "d"])
def test_one_lag_and_zero_lag_equal_born(prec):
    c, w = _case()
    src, rec = c.oracle_sources(), c.oracle_receivers()
    dt = np.float32 if prec == "f" else np.float64
    wav = [w[None].astype(dt)]
    rng = np.random.default_rng(22)
    m = rng.standard_normal((c.nx, c.nz)).astype(dt)
    born = O.born_forward(c.grid, c.vel, src, wav, rec, m, prec)
    g1 = c.grid
    g1.nExt = 1
    np.testing.assert_array_equal(O.borntime_forward(g1, c.vel, src, wav, rec, m[None], 0, prec), born)
    g7 = c.grid
    g7.nExt = 7
    m7 = np.zeros((7, c.nx, c.nz), dt)
    m7[3] = m
    np.testing.assert_array_equal(O.borntime_forward(g7, c.vel, src, wav, rec, m7, 0, prec), born)
    d = rng.standard_normal(born.shape).astype(dt)
    ib = O.born_adjoint(c.grid, c.vel, src, wav, rec, d, prec)
    np.testing.assert_array_equal(O.borntime_adjoint(g1, c.vel, src, wav, rec, d, 0, prec)[0], ib)
    np.testing.assert_array_equal(O.borntime_adjoint(g7, c.vel, src, wav, rec, d, 0, prec)[3], ib)


@pytest.mark.parametrize("lag", [1, 3])
def test_positive_lag_is_born_with_a_delayed_source(lag):
    c, w = _case()
    src, rec = c.oracle_sources(), c.oracle_receivers()
    rng = np.random.default_rng(23)
    m = rng.standard_normal((c.nx, c.nz))
    g = c.grid
    g.nExt = 7
    m7 = np.zeros((7, c.nx, c.nz))
    m7[3 + lag] = m
    wd = np.zeros_like(w)
    wd[2 * lag:] = w[:len(w) - 2 * lag]
    # The last coarse sample of a stored source wavefield only holds the fine
    # steps before it, so the delayed source's W[nts - 1] is not W[nts - 1 - 2 l]:
    # the secondary source differs there, which reaches the last two data
    # samples; the adjoint is compared for data that vanish there.
    born = O.born_forward(c.grid, c.vel, src, [wd[None]], rec, m, "d")
    tl = O.borntime_forward(g, c.vel, src, [w[None]], rec, m7, 0, "d")
    assert np.abs(born).max() > 0
    assert rel_l2(tl[..., :-2], born[..., :-2]) < 1e-12
    # and the adjoint: slice 3 + lag of the time-lag image is the Born image
    # with the delayed source
    d = rng.standard_normal(born.shape)
    d[..., -2:] = 0.0
    ib = O.born_adjoint(c.grid, c.vel, src, [wd[None]], rec, d, "d")
    it = O.borntime_adjoint(g, c.vel, src, [w[None]], rec, d, 0, "d")
    assert rel_l2(it[3 + lag], ib) < 1e-12


def test_sloth_scaling_and_free_surface_ignores_it():
    c, w = _case()
    src, rec = c.oracle_sources(), c.oracle_receivers()
    g = c.grid
    g.nExt = 5
    rng = np.random.default_rng(24)
    m = rng.standard_normal((5, c.nx, c.nz))
    vel = c.vel.astype(np.float64)
    refl = 2.0 / vel ** 3                            # the linearization scale 2/v^3
    a = O.borntime_forward(g, vel, src, [w[None]], rec, m, 1, "d")
    b = O.borntime_forward(g, vel, src, [w[None]], rec, -m / refl[None], 0, "d")
    assert rel_l2(a, b) < 1e-12
    gf = c.grid
    gf.nExt = 5
    gf.freeSurface = 1
    s0 = O.borntime_forward(gf, vel, src, [w[None]], rec, m, 0, "d")
    s1 = O.borntime_forward(gf, vel, src, [w[None]], rec, m, 1, "d")
    np.testing.assert_array_equal(s0, s1)
    assert rel_l2(s0, O.borntime_forward(g, vel, src, [w[None]], rec, m, 0, "d")) > 1e-3


def test_dot_products_and_float_against_fp64():
    c = make_case(seed=25, nShot=2, nts=60, sub=3, dts=0.003)
    g = c.grid
    g.nExt = 9
    src, rec = c.oracle_sources(), c.oracle_receivers()
    rng = np.random.default_rng(26)
    wav = [c.wavelet[None].astype(np.float64)]
    m = rng.standard_normal((9, c.nx, c.nz))
    d = rng.standard_normal((c.nShot, rec[0].nDev, c.nts))
    for sloth in (0, 1):
        fm = O.borntime_forward(g, c.vel, src, wav, rec, m, sloth, "d")
        am = O.borntime_adjoint(g, c.vel, src, wav, rec, d, sloth, "d")
        a, b = np.vdot(fm.ravel(), d.ravel()), np.vdot(m.ravel(), am.ravel())
        print(f"sloth={sloth}: <Fm, d> = {a:.12e}, <m, F'd> = {b:.12e}")
        assert abs(a - b) <= 1e-12 * abs(b)
        ff = O.borntime_forward(g, c.vel, src, [c.wavelet[None]], rec, m.astype(np.float32), sloth, "f")
        af = O.borntime_adjoint(g, c.vel, src, [c.wavelet[None]], rec, d.astype(np.float32), sloth, "f")
        assert rel_l2(ff, fm) < 1e-5
        assert rel_l2(af, am) < 1e-5

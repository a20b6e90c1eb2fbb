"""Wavefield saving on the GPU (forwardWavefield / adjointWavefield and the
getters of PY/pyAcoustic_iso_float_born.cpp:25-33 and
PY/pyAcoustic_iso_float_nl.cpp:34-42) against the float oracle, bitwise:
the nonlinear forward and adjoint wavefields, the Born source wavefield and
its scattered (forward) or scaled receiver (adjoint) wavefield, for shot
wavefieldShotNumber; and the operator outputs are those of forward/adjoint."""
import numpy as np
import pytest

from cases import make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import pyAcoustic_iso_float_born as born
    import pyAcoustic_iso_float_nl as nl
    return nl, born


@pytest.mark.parametrize("sub", [2, 3])
def test_nl_wavefields(mods, sub):
    nl, _ = mods
    c = make_case(seed=71, nShot=3, sub=sub, dts=0.001 * sub, nts=40)
    c.par.set("wavefieldShotNumber", 1)
    op = nl.nonlinearPropShotsGpu(c.vel_vector(), c.par, c.product_sources(nl), c.product_receivers(nl))
    g, os_, or_ = c.grid, c.oracle_sources(), c.oracle_receivers()
    m = c.wavelet[None, None, :].astype(np.float32)
    d = np.zeros((c.nShot, or_[0].nDev, c.nts), np.float32)
    op.forwardWavefield(False, m, d)
    np.testing.assert_array_equal(d, O.nl_forward(g, c.vel, os_, or_, m, "f"))
    wf = op.getWfld()
    ref = O.nl_forward_wavefield(g, c.vel, os_, or_, m, 1)
    print(f"NL forward wavefield sub {sub}: max {np.abs(ref).max():.3e}, bitwise {np.array_equal(wf, ref)}")
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(wf, ref)
    dr = np.random.default_rng(72).standard_normal(d.shape).astype(np.float32)
    ma = np.zeros_like(m)
    op.adjointWavefield(False, ma, dr)
    np.testing.assert_array_equal(ma, O.nl_adjoint(g, c.vel, os_, or_, dr, 1, "f"))
    np.testing.assert_array_equal(op.getWfld(), O.nl_adjoint_wavefield(g, c.vel, os_, or_, dr, 1))


@pytest.mark.parametrize("sub", [2, 3])
def test_born_wavefields(mods, sub):
    nl, born = mods
    c = make_case(seed=73, nShot=3, sub=sub, dts=0.001 * sub, nts=40)
    c.par.set("wavefieldShotNumber", 2)
    wav = c.wavelet[None, :].copy()
    op = born.BornShotsGpu(c.vel_vector(), c.par, c.product_sources(nl), [wav], c.product_receivers(nl))
    g, os_, or_ = c.grid, c.oracle_sources(), c.oracle_receivers()
    rng = np.random.default_rng(74)
    m = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    d = np.zeros((c.nShot, or_[0].nDev, c.nts), np.float32)
    op.forwardWavefield(False, m, d)
    np.testing.assert_array_equal(d, O.born_forward(g, c.vel, os_, [wav], or_, m, "f"))
    W, Wsec = O.born_wavefields(g, c.vel, os_, [wav], or_, m, 2, adjoint=False)
    np.testing.assert_array_equal(op.getSrcWfld(), W)
    np.testing.assert_array_equal(op.getSecWfld(), Wsec)
    assert np.abs(Wsec).max() > 0
    dr = rng.standard_normal(d.shape).astype(np.float32)
    ma = np.zeros_like(m)
    op.adjointWavefield(False, ma, dr)
    np.testing.assert_array_equal(ma, O.born_adjoint(g, c.vel, os_, [wav], or_, dr, "f"))
    W, Wsec = O.born_wavefields(g, c.vel, os_, [wav], or_, dr, 2, adjoint=True)
    np.testing.assert_array_equal(op.getSrcWfld(), W)
    np.testing.assert_array_equal(op.getSecWfld(), Wsec)
    assert np.abs(Wsec).max() > 0


def test_python_layer_wavefield_vectors(tmp_path):
    """The Python operator classes return the saved wavefields as vectors on
    the velocity's z and x axes and the time axis."""
    from acoustic_isotropic_operators_amd import Acoustic_iso_float as AI
    from test_operator_layer import _write_case_files
    c = make_case(seed=75, nShot=2, sub=2, dts=0.002, nts=30)
    args = _write_case_files(tmp_path, c) + ["wavefieldShotNumber=1"]
    m, d, vel, par, srcs, sig, recs, _ = AI.BornOpInitFloat(args)
    op = AI.BornShotsGpu(m, d, vel, par, srcs, sig, recs)
    m.getNdArray()[...] = np.random.default_rng(76).standard_normal(m.getNdArray().shape)
    op.forwardWavefield(False, m, d)
    w = op.getSrcWfld()
    ax = w.getHyper().axes
    assert w.getNdArray().shape == (c.nts, c.nx, c.nz)
    assert (ax[0].n, ax[1].n, ax[2].n) == (c.nz, c.nx, c.nts) and ax[2].d == pytest.approx(c.dts)
    assert np.abs(op.getSecWfld().getNdArray()).max() > 0


def test_getter_before_saving_raises(mods):
    nl, born = mods
    c = make_case(seed=77, nShot=1)
    op = nl.nonlinearPropShotsGpu(c.vel_vector(), c.par, c.product_sources(nl), c.product_receivers(nl))
    with pytest.raises(RuntimeError, match="no wavefield saved"):
        op.getWfld()

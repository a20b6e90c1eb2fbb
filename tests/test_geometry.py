"""Acquisition geometry (deviceGpu) of the product's host code against the
oracle and against the integer closed forms -- bit-exact.  CPU only."""
import numpy as np
import pytest

from cases import make_case
from oracle import oracle as O


@pytest.fixture(scope="module")
def nl():
    import pyAcoustic_iso_float_nl as m
    return m


@pytest.mark.parametrize("kind,per_shot", [("grid", False), ("coords", True), ("ints", True)])
def test_product_geometry_matches_oracle(nl, kind, per_shot):
    c = make_case(seed=1, nShot=3, rec_kind=kind, per_shot_rec=per_shot, src_kind="coords" if per_shot else "grid")
    for o, p in zip(c.oracle_receivers() + c.oracle_sources(), c.product_receivers(nl) + c.product_sources(nl)):
        assert p.getNDeviceIrreg() == o.nDev
        assert p.getNDeviceReg() == o.nReg
        np.testing.assert_array_equal(np.array(p.getRegPosUnique()), o.regPos)
        np.testing.assert_array_equal(np.array(p.getRegPos()), o.gridIdx)
        np.testing.assert_array_equal(np.array(p.getWeights(), np.float32), np.asarray(o.weight, np.float32))


def test_grid_receivers_closed_form(nl):
    """Constructor #3 on a horizontal line: corners b, b+1, b+nz, b+nz+1 with
    first-seen dedup give 2*(nRec+1) points in the order
    b0, b0+1, b0+nz, b0+nz+1, b1+nz, b1+nz+1, ... (F/src/deviceGpu.cpp:142-224)."""
    nz, nx, nrec, iz, ix0 = 618, 1322, 1200, 57, 55
    v = np.ones((nx, nz), np.float32)
    d = nl.deviceGpu(1, iz, 1, nrec, ix0, 1, v, 10)
    pos = np.array(d.getRegPosUnique())
    b = np.array([(ix0 + i) * nz + iz for i in range(nrec)])
    expect = [b[0], b[0] + 1] + [x for i in range(nrec) for x in (b[i] + nz, b[i] + nz + 1)]
    assert d.getNDeviceReg() == 2 * (nrec + 1)
    np.testing.assert_array_equal(pos, np.array(expect))
    o = O.Device.grid(1, iz, 1, nrec, ix0, 1, nz, nx, 10)
    np.testing.assert_array_equal(o.regPos, np.array(expect))


def test_out_of_bounds_raises(nl):
    v = np.ones((42, 26), np.float32)
    with pytest.raises(RuntimeError, match="out of bounds"):
        nl.deviceGpu(1, 30, 1, 1, 8, 1, v, 8)
    with pytest.raises(RuntimeError, match="out of bounds"):
        nl.deviceGpu([3, 30], [4, 4], v, 8)
    with pytest.raises(RuntimeError, match="out of bounds"):
        nl.deviceGpu(np.array([250.0], np.float32), np.array([1.0], np.float32), v, 8)


def test_device_forward_adjoint_match_oracle(nl):
    c = make_case(seed=3, nShot=1, rec_kind="coords", per_shot_rec=True)
    o = c.oracle_receivers()[0]
    p = c.product_receivers(nl)[0]
    rng = np.random.default_rng(0)
    reg = rng.standard_normal((o.nReg, c.nts)).astype(np.float32)
    irr = np.zeros((o.nDev, c.nts), np.float32)
    p.forward(False, reg, irr)
    np.testing.assert_array_equal(irr, o.forward(reg))
    d = rng.standard_normal((o.nDev, c.nts)).astype(np.float32)
    r2 = np.zeros_like(reg)
    p.adjoint(False, r2, d)
    np.testing.assert_array_equal(r2, o.adjoint(d))

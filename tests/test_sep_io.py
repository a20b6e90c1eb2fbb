"""SEP .H/.H@ input and output (sep.read / sep.write) on the CPU: round
trips in both byte orders, headers with history blocks (the last in= wins,
n3..n9 = 1 are not dimensions), size and format checks, and the
reference's own data file notebooks/velocityMarmousi.H when it is present
in this container."""
import os

import numpy as np
import pytest

from acoustic_isotropic_operators_amd import sep

MARMOUSI = "/root/reference/acoustic_iso_lib/notebooks/velocityMarmousi.H"


@pytest.mark.parametrize("fmt", ["xdr_float", "native_float"])
def test_round_trip(tmp_path, fmt):
    a = np.random.default_rng(1).standard_normal((3, 5, 7)).astype(np.float32)
    v = sep.FloatVector(sep.hypercube(ns=[7, 5, 3], os_=[0.5, -1.0, 2.0], ds=[0.1, 0.2, 0.3]), a)
    sep.write(str(tmp_path / "a.H"), v, fmt=fmt)
    raw = np.fromfile(str(tmp_path / "a.H@"), dtype=">f4" if fmt == "xdr_float" else "<f4")
    np.testing.assert_array_equal(raw, a.ravel())
    b = sep.read(str(tmp_path / "a.H"))
    np.testing.assert_array_equal(b.getNdArray(), a)
    ax = b.getHyper().axes
    assert [x.n for x in ax] == [7, 5, 3]
    assert ax[1].o == pytest.approx(-1.0) and ax[2].d == pytest.approx(0.3)


def test_history_header(tmp_path):
    """A header written by a chain of SEP programs: the last in= names the
    binary (relative to the header), quoted values, comments, and trailing
    n3..n9 = 1 that do not add dimensions."""
    a = np.arange(6 * 4, dtype=np.float32).reshape(4, 6)
    a.astype(">f4").tofile(str(tmp_path / "v.H@"))
    (tmp_path / "v.H").write_text(
        'Prog1:  user@host\n\t\tsets next: in="/no/such/dir/first.H@"\n'
        '\tn1=9 o1=0 d1=1 label1="z [m]"\n n3=1 n4=1 n5=1 n6=1 n7=1 n8=1 n9=1\n'
        '#Window #n1=6\n'
        'Prog2:  user@host\n\tn1=6  o1=0.000000  d1=0.010000   label1="z [m]"\n'
        '\tn2=4  o2=0.000000  d2=0.020000   label2="x [m]"\n\t\tesize=4\n\t\tdata_format="xdr_float"\n'
        "in='./v.H@'\n")
    v = sep.read(str(tmp_path / "v.H"))
    assert v.getNdArray().shape == (4, 6)
    np.testing.assert_array_equal(v.getNdArray(), a)
    assert sep.read(str(tmp_path / "v.H"), ndims=3).getNdArray().shape == (1, 4, 6)


def test_short_binary_and_unsupported_format(tmp_path):
    np.zeros(5, ">f4").tofile(str(tmp_path / "s.H@"))
    (tmp_path / "s.H").write_text('n1=6 n2=1\ndata_format="xdr_float"\nin="s.H@"\n')
    with pytest.raises(ValueError, match="holds 5 samples"):
        sep.read(str(tmp_path / "s.H"))
    (tmp_path / "c.H").write_text('n1=6\nesize=8\ndata_format="xdr_complex"\nin="s.H@"\n')
    with pytest.raises(ValueError, match="only 4-byte"):
        sep.read(str(tmp_path / "c.H"))


@pytest.mark.skipif(not os.path.exists(MARMOUSI), reason="the reference's notebooks are not in this container")
def test_reference_marmousi_file():
    v = sep.read(MARMOUSI)
    a = v.getNdArray()
    assert a.shape == (1700, 350)                      # n1 = 350 (z), n2 = 1700 (x)
    ax = v.getHyper().axes
    assert ax[0].d == pytest.approx(0.01) and ax[1].d == pytest.approx(0.01)
    assert 1.49 < float(a.min()) < 1.51 and 4.49 < float(a.max()) < 4.51   # km/s
    assert np.all(np.diff(a[:, :40].mean(axis=0)) > -1e-3)    # water layer then increasing with depth at the top

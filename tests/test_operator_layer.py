"""The Python operator layer (Acoustic_iso_float): parfile handling, the
acquisition builders and the *OpInitFloat builders on the CPU; the operator
classes on the GPU against the oracle."""
import numpy as np
import pytest

from acoustic_isotropic_operators_amd import Acoustic_iso_float as AI
from acoustic_isotropic_operators_amd import sep
from cases import make_case, rel_l2
from oracle import oracle as O

FAT = 5


def _write_case_files(tmp_path, c, nExt=1, geom_files=False):
    """Velocity, wavelet and (optionally) geometry files of a test case, and the
    reference parfile keys that describe it."""
    vel = sep.FloatVector(sep.hypercube(ns=[c.nz, c.nx], os_=[0.0, 0.0], ds=[c.dz, c.dx]), c.vel)
    sep.write(str(tmp_path / "vel.H"), vel)
    wav = sep.FloatVector(sep.hypercube(ns=[c.nts, 1], os_=[0.0, 0.0], ds=[c.dts, 1.0]), c.wavelet[None, :])
    sep.write(str(tmp_path / "wav.H"), wav)
    args = [f"vel={tmp_path / 'vel.H'}", f"sources={tmp_path / 'wav.H'}", f"nts={c.nts}", f"dts={c.dts}",
            f"sub={c.sub}", f"nShot={c.nShot}", f"zPadMinus={c.pad}", f"zPadPlus={c.pad}", f"xPadMinus={c.pad}",
            f"xPadPlus={c.pad}", "fat=5", f"alphaCos={c.alpha}", "fMax=20", f"nz={c.nz}", f"nx={c.nx}",
            f"dz={c.dz}", f"dx={c.dx}", "nGpu=1", "blockSize=16"]
    # zSource=12 -> iz = 12-1+pad+fat = 26 etc., in padded-grid points
    args += ["zSource=2", "xSource=1", "spacingShots=12", "depthReceiver=2", "nReceiver=20", "oReceiver=3",
             "dReceiver=2"]
    if nExt > 1:
        args += ["extension=offset", f"nExt={nExt}"]
    if geom_files:
        # The reference's layout (PY/Acoustic_iso_float.py:250-259, 414-431):
        # arrays [3][nShot] and [3][nRec][nShot], i.e. SEP axis 1 = shot.
        src = np.zeros((3, c.nShot), np.float32)
        for s in range(c.nShot):
            src[:, s] = [(FAT + c.pad + 3 + 10 * s) * c.dx + 2.5, 0.0, (FAT + c.pad + 1) * c.dz + 1.5]
        sep.write(str(tmp_path / "srcgeom.H"), sep.FloatVector(sep.hypercube(ns=[c.nShot, 3]), src))
        nrec = 7
        rec = np.zeros((3, nrec, c.nShot), np.float32)
        for s in range(c.nShot):
            for r in range(nrec):
                rec[:, r, s] = [(FAT + c.pad + 2 + 4 * r + s) * c.dx, 0.0, (FAT + c.pad + 2) * c.dz + 4.0]
        sep.write(str(tmp_path / "recgeom.H"), sep.FloatVector(sep.hypercube(ns=[c.nShot, nrec, 3]), rec))
        args += [f"sourceGeomFile={tmp_path / 'srcgeom.H'}", f"receiverGeomFile={tmp_path / 'recgeom.H'}",
                 f"nReceiverPerShot={nrec}"]
    return args


def test_parfile_expansion(tmp_path):
    (tmp_path / "p.par").write_text("nts=100\ndts=0.004 sub=4\n")
    par = AI.create_parObj(["nShot=3", f"par={tmp_path / 'p.par'}", "nts=200"])
    assert par.getInt("nts") == 200 and par.getInt("sub") == 4 and par.getFloat("dts") == pytest.approx(0.004)


def test_regular_geometry_builders(tmp_path):
    c = make_case(nShot=3)
    par = AI.create_parObj(_write_case_files(tmp_path, c))
    vel = sep.read(str(tmp_path / "vel.H"), ndims=2)
    srcs, sax = AI.buildSourceGeometry(par, vel)
    recs, rax = AI.buildReceiversGeometry(par, vel)
    assert len(srcs) == 3 and len(recs) == 1
    izs = 2 - 1 + c.pad + FAT
    for s, d in enumerate(srcs):
        ix = 1 - 1 + c.pad + FAT + 12 * s
        assert d.getRegPosUnique()[0] == ix * c.nz + izs
    assert recs[0].getNDeviceIrreg() == 20
    assert rax.n == 20 and rax.d == pytest.approx(2 * c.dx)
    assert sax.n == 3 and sax.d == pytest.approx(12 * c.dx)


@pytest.mark.parametrize("nShot", [2, 3])
def test_geometry_file_builders(tmp_path, nShot):
    """Geometry files in the reference's layout, shot on SEP axis 1.  With
    nShot == 3 a reader that took the coordinate axis for the shot axis would
    accept the file and misread it, so the positions are checked per shot."""
    c = make_case(nShot=nShot)
    par = AI.create_parObj(_write_case_files(tmp_path, c, geom_files=True))
    vel = sep.read(str(tmp_path / "vel.H"), ndims=2)
    srcs, _ = AI.buildSourceGeometry(par, vel)
    recs, rax = AI.buildReceiversGeometry(par, vel)
    assert len(srcs) == nShot and len(recs) == nShot and rax.n == 7
    gs = sep.read(str(tmp_path / "srcgeom.H"), ndims=2).getNdArray()
    gr = sep.read(str(tmp_path / "recgeom.H"), ndims=3).getNdArray()
    assert gs.shape == (3, nShot) and gr.shape == (3, 7, nShot)
    # bilinear constructor on physical coordinates: the oracle restates it
    for s in range(nShot):
        o = O.Device.coords(gs[2, s:s + 1], gs[0, s:s + 1], c.nz, c.nx, 0.0, c.dz, 0.0, c.dx, c.nts)
        np.testing.assert_array_equal(np.array(srcs[s].getRegPosUnique()), o.regPos)
        o = O.Device.coords(gr[2, :, s], gr[0, :, s], c.nz, c.nx, 0.0, c.dz, 0.0, c.dx, c.nts)
        np.testing.assert_array_equal(np.array(recs[s].getRegPosUnique()), o.regPos)
    # the sources move along x by 10 grid points per shot (src file layout pinned)
    ix = [srcs[s].getRegPosUnique()[0] // c.nz for s in range(nShot)]
    assert np.all(np.diff(ix) == 10)


def test_geometry_file_transposed_is_rejected(tmp_path):
    """A file written with the coordinate axis fastest ([nShot][3]) does not
    match the reference's shape checks and is refused, as by the reference."""
    c = make_case(nShot=2)
    args = _write_case_files(tmp_path, c, geom_files=True)
    src = np.zeros((c.nShot, 3), np.float32)
    src[:, 0] = 300.0
    src[:, 2] = 200.0
    sep.write(str(tmp_path / "srcT.H"), sep.FloatVector(sep.hypercube(ns=[3, c.nShot]), src))
    args = [a for a in args if not a.startswith("sourceGeomFile=")] + [f"sourceGeomFile={tmp_path / 'srcT.H'}"]
    par = AI.create_parObj(args)
    vel = sep.read(str(tmp_path / "vel.H"), ndims=2)
    with pytest.raises(ValueError, match="not consistent with geometry file"):
        AI.buildSourceGeometry(par, vel)


def test_geometry_file_shot_count_is_checked(tmp_path):
    c = make_case(nShot=2)
    args = _write_case_files(tmp_path, c, geom_files=True) + ["nShot=3"]
    par = AI.create_parObj(args)
    vel = sep.read(str(tmp_path / "vel.H"), ndims=2)
    with pytest.raises(ValueError, match="not consistent"):
        AI.buildSourceGeometry(par, vel)


def test_opinit_shapes(tmp_path):
    c = make_case(nShot=2)
    args = _write_case_files(tmp_path, c, nExt=5)
    m, d, vel, par, srcs, sig, recs, _ = AI.BornOpInitFloat(args)
    assert m.getNdArray().shape == (c.nx, c.nz)
    assert d.getNdArray().shape == (2, 20, c.nts)
    assert sig[0].getNdArray().shape == (1, c.nts)
    m, d, *_ = AI.BornExtOpInitFloat(args)
    assert m.getNdArray().shape == (5, c.nx, c.nz)
    ext = m.getHyper().axes[2]
    assert ext.o == pytest.approx(-2 * c.dx) and ext.d == pytest.approx(c.dx)
    m, d, *_ = AI.nonlinearOpInitFloat(args)
    assert m.getNdArray().shape == (1, 1, c.nts)


def test_dask_client_is_refused(tmp_path):
    c = make_case(nShot=1)
    with pytest.raises(NotImplementedError):
        AI.BornOpInitFloat(_write_case_files(tmp_path, c), client=object())


@pytest.mark.gpu
def test_born_operator_class_matches_oracle(tmp_path):
    c = make_case(seed=31, nShot=2, sub=2, dts=0.002)
    args = _write_case_files(tmp_path, c)
    m, d, vel, par, srcs, sig, recs, _ = AI.BornOpInitFloat(args)
    op = AI.BornShotsGpu(m, d, vel, par, srcs, sig, recs)
    rng = np.random.default_rng(32)
    m.getNdArray()[...] = rng.standard_normal(m.getNdArray().shape)
    op.forward(False, m, d)
    g = c.grid
    osrc = [O.Device.grid(1, 2 - 1 + c.pad + FAT, 1, 1, c.pad + FAT + 12 * s, 1, c.nz, c.nx, c.nts) for s in range(2)]
    orec = [O.Device.grid(1, 2 - 1 + c.pad + FAT, 1, 20, 3 - 1 + c.pad + FAT, 2, c.nz, c.nx, c.nts)]
    wav = sig[0].getNdArray()
    ref = O.born_forward(g, c.vel, osrc, [wav], orec, m.getNdArray(), "f")
    np.testing.assert_array_equal(d.getNdArray(), ref)
    assert op.dotTest(verb=True, maxError=1e-4)
    assert op.dotTestCpp(False, 1e-5)


def test_tomo_opinit_shapes_and_reflectivity_file(tmp_path):
    c = make_case(nShot=2)
    args = _write_case_files(tmp_path, c, nExt=5)
    m, d, vel, par, srcs, sig, recs, refl, _ = AI.tomoExtOpInitFloat(args)
    assert m.getNdArray().shape == (c.nx, c.nz)
    assert refl.getNdArray().shape == (5, c.nx, c.nz) and not np.any(refl.getNdArray())
    r = np.random.default_rng(1).standard_normal((5, c.nx, c.nz)).astype(np.float32)
    sep.write(str(tmp_path / "refl.H"), sep.FloatVector(refl.getHyper(), r))
    *_, refl2, _ = AI.tomoExtOpInitFloat(args + [f"reflectivity={tmp_path / 'refl.H'}"])
    np.testing.assert_array_equal(refl2.getNdArray(), r)


@pytest.mark.gpu
def test_tomo_operator_class_matches_oracle(tmp_path):
    c = make_case(seed=35, nShot=2, sub=2, dts=0.002, nExt=5)
    args = _write_case_files(tmp_path, c, nExt=5)
    m, d, vel, par, srcs, sig, recs, refl, _ = AI.tomoExtOpInitFloat(args)
    rng = np.random.default_rng(36)
    refl.getNdArray()[...] = rng.standard_normal(refl.getNdArray().shape)
    op = AI.tomoExtShotsGpu(m, d, vel, par, srcs, sig, recs, refl)
    m.getNdArray()[...] = rng.standard_normal(m.getNdArray().shape)
    op.forward(False, m, d)
    g = c.grid
    osrc = [O.Device.grid(1, 2 - 1 + c.pad + FAT, 1, 1, c.pad + FAT + 12 * s, 1, c.nz, c.nx, c.nts) for s in range(2)]
    orec = [O.Device.grid(1, 2 - 1 + c.pad + FAT, 1, 20, 3 - 1 + c.pad + FAT, 2, c.nz, c.nx, c.nts)]
    wav = sig[0].getNdArray()
    ref = O.tomo_forward(g, c.vel, osrc, [wav], orec, m.getNdArray(), refl.getNdArray())
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(d.getNdArray(), ref)
    dr = d.clone()
    dr.getNdArray()[...] = rng.standard_normal(dr.getNdArray().shape)
    op.adjoint(False, m, dr)
    np.testing.assert_array_equal(m.getNdArray(), O.tomo_adjoint(g, c.vel, osrc, [wav], orec, dr.getNdArray(),
                                                                 refl.getNdArray()))
    # a new reflectivity is taken by the next call
    refl2 = refl.clone()
    refl2.getNdArray()[...] *= 2.0
    op.setReflectivityExt(refl2)
    op.forward(False, m, d)
    np.testing.assert_array_equal(d.getNdArray(), O.tomo_forward(g, c.vel, osrc, [wav], orec, m.getNdArray(),
                                                                 refl2.getNdArray()))
    assert op.dotTestCpp(False, 1e-5)


def _write_wemva_data(tmp_path, c, seed=1):
    dobs = np.random.default_rng(seed).standard_normal((c.nShot, 20, c.nts)).astype(np.float32)
    sep.write(str(tmp_path / "dobs.H"), sep.FloatVector(sep.hypercube(ns=[c.nts, 20, c.nShot]), dobs))
    return dobs, [f"seismicData={tmp_path / 'dobs.H'}"]


def test_wemva_opinit_shapes(tmp_path):
    c = make_case(nShot=2)
    args = _write_case_files(tmp_path, c, nExt=5)
    dobs, extra = _write_wemva_data(tmp_path, c)
    m, d, vel, par, srcs, sig, recs, rsig = AI.wemvaExtOpInitFloat(args + extra)
    assert m.getNdArray().shape == (c.nx, c.nz)
    assert d.getNdArray().shape == (5, c.nx, c.nz)
    assert len(rsig) == 2
    np.testing.assert_array_equal(np.stack([r.getNdArray() for r in rsig]), dobs)
    with pytest.raises(ValueError, match="seismic data"):
        AI.wemvaExtOpInitFloat(args)


@pytest.mark.gpu
def test_wemva_operator_class_matches_oracle(tmp_path):
    c = make_case(seed=37, nShot=2, sub=2, dts=0.002, nExt=5)
    args = _write_case_files(tmp_path, c, nExt=5)
    dobs, extra = _write_wemva_data(tmp_path, c, seed=38)
    m, im, vel, par, srcs, sig, recs, rsig = AI.wemvaExtOpInitFloat(args + extra)
    op = AI.wemvaExtShotsGpu(m, im, vel, par, srcs, sig, recs, rsig)
    rng = np.random.default_rng(39)
    m.getNdArray()[...] = rng.standard_normal(m.getNdArray().shape)
    op.forward(False, m, im)
    g = c.grid
    osrc = [O.Device.grid(1, 2 - 1 + c.pad + FAT, 1, 1, c.pad + FAT + 12 * s, 1, c.nz, c.nx, c.nts) for s in range(2)]
    orec = [O.Device.grid(1, 2 - 1 + c.pad + FAT, 1, 20, 3 - 1 + c.pad + FAT, 2, c.nz, c.nx, c.nts)]
    wav = sig[0].getNdArray()
    ref = O.wemva_forward(g, c.vel, osrc, [wav], orec, dobs, m.getNdArray())
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(im.getNdArray(), ref)
    im.getNdArray()[...] = rng.standard_normal(im.getNdArray().shape)
    op.adjoint(False, m, im)
    np.testing.assert_array_equal(m.getNdArray(), O.wemva_adjoint(g, c.vel, osrc, [wav], orec, dobs, im.getNdArray()))
    # Float dot test: the extended image of random vectors sums many terms of
    # both signs, and the float relative error is 5e-5 here (the FP64 oracle's
    # dot tests hold to 1e-12); the engine is bitwise equal to the float oracle.
    assert op.dotTestCpp(False, 1e-4)

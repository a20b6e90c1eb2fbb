"""Time-lag extended Born (extension=time) on the GPU against the float
oracle, bit for bit: batched lag imaging and secondary sources for every
batch size, the window-capture path (sub = 2) and the accumulate path
(sub = 1, 3), lags that reach past both ends of the time axis, the sloth
scaling, a free surface, and the pybind operator."""
import numpy as np
import pytest

from cases import make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _engine(c, fs=0):
    from acoustic_isotropic_operators_amd import _capi as A
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub,
                  zPadMinus=0 if fs else c.pad, zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad,
                  alphaCos=c.alpha, fMax=20.0, nExt=c.nExt, extension=A.EXT_TIME, freeSurface=fs)
    ctx = A.Context(params)
    ctx.set_velocity(c.vel)
    srcs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.src_args]
    recs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.rec_args]
    return ctx, srcs, recs


def _check(c, g, ctx, srcs, recs, sloth, batches, seed):
    os_, or_ = c.oracle_sources(), c.oracle_receivers()
    wav = c.wavelet[None, :].astype(np.float32)
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((c.nExt, c.nx, c.nz)).astype(np.float32)
    dr = rng.standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)
    of = O.borntime_forward(g, c.vel, os_, [wav], or_, m, sloth, "f")
    af = O.borntime_adjoint(g, c.vel, os_, [wav], or_, dr, sloth, "f")
    assert np.abs(of).max() > 0 and np.abs(af).max() > 0
    ctx.set_option("sloth", sloth)
    for k in batches:
        print(f"nts {c.nts} sub {c.sub} nExt {c.nExt} fs {g.freeSurface} sloth {sloth} extBatch {k}", flush=True)
        ctx.set_option("extBatch", k)
        d = np.zeros_like(dr)
        ctx.bornext_forward(m, d, srcs, wav, recs)
        np.testing.assert_array_equal(d, of, err_msg=f"forward, extBatch {k}")
        ma = np.zeros_like(m)
        ctx.bornext_adjoint(ma, dr, srcs, wav, recs)
        np.testing.assert_array_equal(ma, af, err_msg=f"adjoint, extBatch {k}")


@pytest.mark.parametrize("sub,nExt,sloth", [(2, 11, 0), (3, 7, 1), (1, 31, 0), (2, 1, 0)])
def test_time_lag_matches_oracle(sub, nExt, sloth):
    # 23 coarse samples: batches of 5 and 8 leave a partial last batch; 31
    # lags (hExt = 15) reach past both ends of the time axis
    c = make_case(seed=71, nShot=3, nts=23, sub=sub, dts=0.001 * sub, nExt=nExt)
    ctx, srcs, recs = _engine(c)
    _check(c, c.grid, ctx, srcs, recs, sloth, (1, 5, 8, 16), 72)


@pytest.mark.parametrize("nExt,nts", [(5, 80), (13, 90), (41, 120)])
def test_full_interior_batches_match_oracle(nExt, nts):
    # long enough for batches of 16 samples with every lag usable: the
    # register-window kernels (hExt up to 5, 10, 20) next to the general ones
    c = make_case(seed=77, nShot=2, nts=nts, sub=1, dts=0.001, nExt=nExt)
    ctx, srcs, recs = _engine(c)
    _check(c, c.grid, ctx, srcs, recs, 0, (16, 7), 78)


def test_time_lag_free_surface_matches_oracle():
    c = make_case(seed=73, nShot=2, nts=30, sub=2, dts=0.002, nExt=9)
    ctx, srcs, recs = _engine(c, fs=1)
    g = c.grid
    g.freeSurface = 1
    _check(c, g, ctx, srcs, recs, 1, (16,), 74)     # sloth is ignored with a free surface


def test_pybind_operator():
    import pyAcoustic_iso_float_born_ext as bx
    import pyAcoustic_iso_float_nl as nl
    c = make_case(seed=75, nShot=2, nts=40, sub=2, dts=0.002, nExt=9)
    c.par.set("extension", "time")
    op = bx.BornExtShotsGpu(c.vel_vector(), c.par, c.product_sources(nl), [c.wavelet[None, :].copy()],
                            c.product_receivers(nl))
    rng = np.random.default_rng(76)
    m = rng.standard_normal((c.nExt, c.nx, c.nz)).astype(np.float32)
    d = np.zeros((c.nShot, c.oracle_receivers()[0].nDev, c.nts), np.float32)
    op.forward(False, m, d)
    os_, or_ = c.oracle_sources(), c.oracle_receivers()
    np.testing.assert_array_equal(d, O.borntime_forward(c.grid, c.vel, os_, [c.wavelet[None, :]], or_, m, 0, "f"))
    dr = rng.standard_normal(d.shape).astype(np.float32)
    img = np.zeros_like(m)
    op.adjoint(False, img, dr)
    np.testing.assert_array_equal(img, O.borntime_adjoint(c.grid, c.vel, os_, [c.wavelet[None, :]], or_, dr, 0, "f"))
    assert op.dotTest(True, 1e-5)

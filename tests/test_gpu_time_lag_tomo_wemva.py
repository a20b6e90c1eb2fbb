"""Time-lag tomographic and WEMVA operators (extension=time) on the GPU
through the C-ABI, bitwise equal to the float oracle: each leg and both, the
window-capture (sub = 2) and accumulate (sub = 3) source passes, ring
batches smaller than the number of coarse samples, lags past both ends of
the time axis, the sloth scaling, and a free surface (which the WEMVA
operator ignores)."""
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


def _run(c, g, ctx, srcs, recs, sloth, extBatch, seed, legs_list=((1, 1), (1, 0), (0, 1))):
    os_, or_ = c.oracle_sources(), c.oracle_receivers()
    wav = c.wavelet[None, :].astype(np.float32)
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((c.nExt, c.nx, c.nz)).astype(np.float32)
    m = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    dr = rng.standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)
    ctx.set_option("extBatch", extBatch)
    ctx.set_option("sloth", sloth)
    ctx.set_ext_reflectivity(r)
    for legs in legs_list:
        print(f"nts {c.nts} sub {c.sub} nExt {c.nExt} fs {g.freeSurface} sloth {sloth} extBatch {extBatch} "
              f"legs {legs}", flush=True)
        ctx.set_option("leg1", legs[0])
        ctx.set_option("leg2", legs[1])
        of = O.tomo_forward(g, c.vel, os_, [wav], or_, m, r, *legs, time=True, sloth=sloth)
        d = np.zeros_like(dr)
        ctx.tomo_forward(m, d, srcs, wav, recs)
        np.testing.assert_array_equal(d, of, err_msg=f"tomo forward, legs {legs}")
        af = O.tomo_adjoint(g, c.vel, os_, [wav], or_, dr, r, *legs, time=True, sloth=sloth)
        ma = np.zeros_like(m)
        ctx.tomo_adjoint(ma, dr, srcs, wav, recs)
        np.testing.assert_array_equal(ma, af, err_msg=f"tomo adjoint, legs {legs}")
        wf = O.wemva_forward(g, c.vel, os_, [wav], or_, dr, m, *legs, time=True)
        im = np.zeros_like(r)
        ctx.wemva_forward(m, im, dr, srcs, wav, recs)
        np.testing.assert_array_equal(im, wf, err_msg=f"WEMVA forward, legs {legs}")
        wa = O.wemva_adjoint(g, c.vel, os_, [wav], or_, dr, r, *legs, time=True)
        mw = np.zeros_like(m)
        ctx.wemva_adjoint(mw, r, dr, srcs, wav, recs)
        np.testing.assert_array_equal(mw, wa, err_msg=f"WEMVA adjoint, legs {legs}")
        assert min(np.abs(x).max() for x in (of, af, wf, wa)) > 0


@pytest.mark.parametrize("sub,nExt,sloth,extBatch", [(2, 7, 0, 16), (3, 5, 1, 4), (2, 31, 0, 5)])
def test_time_lag_tomo_and_wemva_match_oracle(sub, nExt, sloth, extBatch):
    # 23 coarse samples; 31 lags (hExt = 15) reach past both ends
    c = make_case(seed=91, nShot=3, nts=23, sub=sub, dts=0.001 * sub, nExt=nExt)
    ctx, srcs, recs = _engine(c)
    _run(c, c.grid, ctx, srcs, recs, sloth, extBatch, 92)


@pytest.mark.parametrize("nExt,sloth", [(9, 0), (21, 1)])
def test_full_interior_batches_match_oracle(nExt, sloth):
    # batches of 16 samples with every lag usable in both lag directions and
    # both sample orders (the register-window kernels)
    c = make_case(seed=95, nShot=2, nts=100, sub=1, dts=0.001, nExt=nExt)
    ctx, srcs, recs = _engine(c)
    _run(c, c.grid, ctx, srcs, recs, sloth, 16, 96)


def test_time_lag_tomo_and_wemva_with_free_surface():
    c = make_case(seed=93, nShot=2, nts=30, sub=2, dts=0.002, nExt=7)
    ctx, srcs, recs = _engine(c, fs=1)
    g = c.grid
    g.freeSurface = 1
    _run(c, g, ctx, srcs, recs, 1, 16, 94, legs_list=((1, 1),))

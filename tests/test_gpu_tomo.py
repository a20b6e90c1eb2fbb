"""Offset-extended tomographic operator on the GPU through the C-ABI
(aiso_tomo_forward / aiso_tomo_adjoint) against the float oracle: bitwise
equal for each leg and both, on the window-capture (sub = 2) and the
accumulate (sub = 3) source pass, and for ring batches smaller than the
number of coarse samples."""
import numpy as np
import pytest

from cases import make_case, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _engine(c):
    from acoustic_isotropic_operators_amd import _capi as A
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=c.pad,
                  zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0, nExt=c.nExt,
                  extension=A.EXT_OFFSET)
    ctx = A.Context(params)
    ctx.set_velocity(c.vel)
    srcs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.src_args]
    recs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.rec_args]
    return ctx, srcs, recs


# nExt 3/5/11/21 reach the tiled adjoint offset scattering for hExt <= 2, 5
# and 10; nExt 23 (hExt 11) the untiled kernel.
@pytest.mark.parametrize("sub,nExt,extBatch", [(2, 5, 16), (3, 3, 4), (2, 21, 5), (2, 11, 16), (2, 23, 8)])
def test_tomo_matches_oracle(sub, nExt, extBatch):
    c = make_case(seed=81, nShot=3, nts=23, sub=sub, dts=0.001 * sub, nExt=nExt)
    ctx, srcs, recs = _engine(c)
    ctx.set_option("extBatch", extBatch)
    g, os_, or_ = c.grid, c.oracle_sources(), c.oracle_receivers()
    wav = c.wavelet[None, :].astype(np.float32)
    rng = np.random.default_rng(82)
    r = rng.standard_normal((nExt, c.nx, c.nz)).astype(np.float32)
    m = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    dr = rng.standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)
    ctx.set_ext_reflectivity(r)
    for legs in [(1, 1), (1, 0), (0, 1)]:
        ctx.set_option("leg1", legs[0])
        ctx.set_option("leg2", legs[1])
        of = O.tomo_forward(g, c.vel, os_, [wav], or_, m, r, *legs)
        af = O.tomo_adjoint(g, c.vel, os_, [wav], or_, dr, r, *legs)
        d = np.zeros_like(dr)
        ctx.tomo_forward(m, d, srcs, wav, recs)
        ma = np.zeros_like(m)
        ctx.tomo_adjoint(ma, dr, srcs, wav, recs)
        print(f"sub {sub} nExt {nExt} legs {legs}: fwd relL2 {rel_l2(d, of):.2e} adj relL2 {rel_l2(ma, af):.2e}",
              flush=True)
        assert np.abs(of).max() > 0 and np.abs(af).max() > 0
        np.testing.assert_array_equal(d, of)
        np.testing.assert_array_equal(ma, af)


def test_tomo_add_and_dot_product():
    c = make_case(seed=83, nShot=2, nts=30, sub=2, dts=0.002, nExt=5)
    ctx, srcs, recs = _engine(c)
    rng = np.random.default_rng(84)
    wav = c.wavelet[None, :].astype(np.float32)
    ctx.set_ext_reflectivity(rng.standard_normal((5, c.nx, c.nz)).astype(np.float32))
    m = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    d0 = rng.standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)
    d = d0.copy()
    ctx.tomo_forward(m, d, srcs, wav, recs, add=True)
    dz = np.zeros_like(d0)
    ctx.tomo_forward(m, dz, srcs, wav, recs)
    np.testing.assert_array_equal(d, d0 + dz)
    ma = np.zeros_like(m)
    ctx.tomo_adjoint(ma, d0, srcs, wav, recs)
    a = float(np.vdot(dz.astype(np.float64).ravel(), d0.astype(np.float64).ravel()))
    b = float(np.vdot(m.astype(np.float64).ravel(), ma.astype(np.float64).ravel()))
    print(f"<Fm,d> {a:.8e} <m,F'd> {b:.8e} rel {abs(a - b) / abs(b):.2e}")
    assert abs(a - b) <= 1e-5 * abs(b)


@pytest.mark.parametrize("sub,nExt,extBatch", [(2, 5, 16), (3, 3, 4), (2, 21, 5), (2, 43, 3), (2, 31, 7)])
def test_wemva_matches_oracle(sub, nExt, extBatch):
    """WEMVA (aiso_wemva_forward / aiso_wemva_adjoint), both legs and each
    alone, bitwise equal to the oracle; nExt = 21 and 31 image in offset
    groups of 7 and 14, nExt = 43 takes the untiled imaging kernel."""
    c = make_case(seed=85, nShot=3, nts=23, sub=sub, dts=0.001 * sub, nExt=nExt)
    ctx, srcs, recs = _engine(c)
    ctx.set_option("extBatch", extBatch)
    g, os_, or_ = c.grid, c.oracle_sources(), c.oracle_receivers()
    wav = c.wavelet[None, :].astype(np.float32)
    rng = np.random.default_rng(86)
    dobs = rng.standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)
    m = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    im = rng.standard_normal((nExt, c.nx, c.nz)).astype(np.float32)
    for legs in [(1, 1), (1, 0), (0, 1)]:
        ctx.set_option("leg1", legs[0])
        ctx.set_option("leg2", legs[1])
        of = O.wemva_forward(g, c.vel, os_, [wav], or_, dobs, m, *legs)
        af = O.wemva_adjoint(g, c.vel, os_, [wav], or_, dobs, im, *legs)
        out = np.zeros_like(im)
        ctx.wemva_forward(m, out, dobs, srcs, wav, recs)
        ma = np.zeros_like(m)
        ctx.wemva_adjoint(ma, im, dobs, srcs, wav, recs)
        print(f"sub {sub} nExt {nExt} legs {legs}: fwd relL2 {rel_l2(out, of):.2e} adj relL2 {rel_l2(ma, af):.2e}",
              flush=True)
        assert np.abs(of).max() > 0 and np.abs(af).max() > 0
        np.testing.assert_array_equal(out, of)
        np.testing.assert_array_equal(ma, af)
    ctx.set_option("leg1", 1)
    ctx.set_option("leg2", 1)
    out2 = np.ones_like(im)
    ctx.wemva_forward(m, out2, dobs, srcs, wav, recs, add=True)
    np.testing.assert_array_equal(out2, np.ones_like(im) + O.wemva_forward(g, c.vel, os_, [wav], or_, dobs, m))

"""Free surface (freeSurface = 1, zPadMinus = 0) on the GPU through the
C-ABI: nonlinear, Born, extended-Born, tomographic and WEMVA operators,
bitwise equal to the float oracle, on sub = 2 (window capture, which reads
the stored surface values; the Born forward's lockstep pair) and sub = 3,
with receivers below and on the surface row.  The nonlinear and Born
operators run the surface instances of the specialized kernels, the others
the generic kernel.  The WEMVA operator, like the reference's, propagates with
the absorbing top whatever the key."""
import numpy as np
import pytest

from cases import make_case, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FAT = 5


def _setup(sub, nExt=1, nShot=2, seed=101, nx=74):
    c = make_case(seed=seed, nx=nx, nShot=nShot, nts=30, sub=sub, dts=0.001 * sub, nExt=nExt, src_iz=8)
    from acoustic_isotropic_operators_amd import _capi as A
    zPadPlus = c.pad
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=0,
                  zPadPlus=zPadPlus, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0, nExt=nExt,
                  extension=A.EXT_OFFSET if nExt > 1 else A.EXT_NONE, freeSurface=1)
    ctx = A.Context(params)
    ctx.set_velocity(c.vel)
    g = O.Grid(nz=c.nz, nx=c.nx, nts=c.nts, sub=c.sub, minPad=min(zPadPlus, c.pad), dz=c.dz, dx=c.dx, dts=c.dts,
               alphaCos=c.alpha, nExt=nExt, freeSurface=1)
    src_args = [a for _, a in c.src_args]
    rec_ints = ([FAT, FAT + 1, FAT + 3, 12, 20, 30], [9, 14, 20, 25, 33, 41])   # incl. the surface row
    srcs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for a in src_args]
    recs = [A.Device.ints(*rec_ints, c.nz, c.nx, c.nts)]
    osrc = [O.Device.grid(*a, c.nz, c.nx, c.nts) for a in src_args]
    orec = [O.Device.ints(*rec_ints, c.nz, c.nx, c.nts)]
    return A, c, ctx, g, srcs, recs, osrc, orec


# nx 138: the top row of tiles holds tiles clear of the sponge on every side,
# which the surface kernels run without sponge factors
@pytest.mark.parametrize("sub,nx", [(2, 74), (3, 74), (2, 138), (3, 138)])
def test_nonlinear_and_born_with_free_surface(sub, nx):
    A, c, ctx, g, srcs, recs, osrc, orec = _setup(sub, nx=nx)
    rng = np.random.default_rng(102)
    nrec = recs[0].nDev
    w = c.wavelet[None, None, :].astype(np.float32)
    d = np.zeros((c.nShot, nrec, c.nts), np.float32)
    ctx.nl_forward(w, d, srcs, recs)
    ref = O.nl_forward(g, c.vel, osrc, orec, w, "f")
    print(f"sub {sub}: NL fwd |d| {np.abs(ref).max():.3e}, relL2 {rel_l2(d, ref):.2e}", flush=True)
    assert np.abs(ref).max() > 0 and np.all(ref[:, 0] == 0)          # zero pressure on the surface
    np.testing.assert_array_equal(d, ref)
    dr = rng.standard_normal(d.shape).astype(np.float32)
    ma = np.zeros((c.nShot, 1, c.nts), np.float32)
    ctx.nl_adjoint(ma, dr, srcs, recs)
    np.testing.assert_array_equal(ma, O.nl_adjoint(g, c.vel, osrc, orec, dr, c.nShot, "f"))
    wav = c.wavelet[None, :].astype(np.float32)
    m = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    db = np.zeros_like(d)
    ctx.born_forward(m, db, srcs, wav, recs)
    np.testing.assert_array_equal(db, O.born_forward(g, c.vel, osrc, [wav], orec, m, "f"))
    img = np.zeros_like(m)
    ctx.born_adjoint(img, dr, srcs, wav, recs)
    ref_img = O.born_adjoint(g, c.vel, osrc, [wav], orec, dr, "f")
    print(f"sub {sub}: Born adj |m| {np.abs(ref_img).max():.3e}, relL2 {rel_l2(img, ref_img):.2e}", flush=True)
    np.testing.assert_array_equal(img, ref_img)


def test_extended_born_tomo_and_wemva_with_free_surface():
    A, c, ctx, g, srcs, recs, osrc, orec = _setup(2, nExt=5)
    rng = np.random.default_rng(103)
    wav = c.wavelet[None, :].astype(np.float32)
    nrec = recs[0].nDev
    me = rng.standard_normal((5, c.nx, c.nz)).astype(np.float32)
    d = np.zeros((c.nShot, nrec, c.nts), np.float32)
    ctx.bornext_forward(me, d, srcs, wav, recs)
    np.testing.assert_array_equal(d, O.bornext_forward(g, c.vel, osrc, [wav], orec, me, "f"))
    dr = rng.standard_normal(d.shape).astype(np.float32)
    ie = np.zeros_like(me)
    ctx.bornext_adjoint(ie, dr, srcs, wav, recs)
    np.testing.assert_array_equal(ie, O.bornext_adjoint(g, c.vel, osrc, [wav], orec, dr, "f"))
    ctx.set_ext_reflectivity(me)
    m = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    dt = np.zeros_like(d)
    ctx.tomo_forward(m, dt, srcs, wav, recs)
    np.testing.assert_array_equal(dt, O.tomo_forward(g, c.vel, osrc, [wav], orec, m, me))
    ma = np.zeros_like(m)
    ctx.tomo_adjoint(ma, dr, srcs, wav, recs)
    np.testing.assert_array_equal(ma, O.tomo_adjoint(g, c.vel, osrc, [wav], orec, dr, me))
    im = np.zeros_like(me)
    ctx.wemva_forward(m, im, dr, srcs, wav, recs)
    np.testing.assert_array_equal(im, O.wemva_forward(g, c.vel, osrc, [wav], orec, dr, m))
    mw = np.zeros_like(m)
    ctx.wemva_adjoint(mw, me, dr, srcs, wav, recs)
    np.testing.assert_array_equal(mw, O.wemva_adjoint(g, c.vel, osrc, [wav], orec, dr, me))


def test_free_surface_parameter_checks():
    """fdParam's consistency rule (F/src/fdParam.cpp:41-45) both ways, and
    devices above the surface row are refused."""
    from acoustic_isotropic_operators_amd import _capi as A
    c = make_case(nShot=1)
    base = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadPlus=c.pad,
                xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0)
    with pytest.raises(RuntimeError, match="inconsistent"):
        A.Context(dict(base, zPadMinus=c.pad, freeSurface=1))
    with pytest.raises(RuntimeError, match="inconsistent"):
        A.Context(dict(base, zPadMinus=0, freeSurface=0))
    ctx = A.Context(dict(base, zPadMinus=0, freeSurface=1))
    ctx.set_velocity(c.vel)
    s = [A.Device.grid(1, 3, 1, 1, 20, 1, c.nz, c.nx, c.nts)]
    r = [A.Device.grid(1, 8, 1, 4, 10, 2, c.nz, c.nx, c.nts)]
    with pytest.raises(RuntimeError, match="at or below the surface"):
        ctx.nl_forward(c.wavelet[None, None, :].astype(np.float32), np.zeros((1, 4, c.nts), np.float32), s, r)
    assert A.pad_plus_free_surface(500, 50, 16) == 60        # ceil(550 / 16) * 16 - 500

"""The launch variants of the specialized step kernels give the oracle's
results bit for bit: one or two lanes (a batch's shots split over concurrent
streams) and the register budgets (occupancy 0: the compiler's; 6 or 8 waves
per SIMD), for the nonlinear and Born operators, forward and adjoint,
with sub 2 (window capture, lockstep Born forward) and sub 3 (per-step
accumulation)."""
import numpy as np
import pytest

from cases import make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sub", [2, 3])
@pytest.mark.parametrize("lanes,occ", [(1, 0), (2, 0), (1, 6), (2, 6), (2, 8)])
def test_nl_and_born_bitwise_for_every_variant(lanes, occ, sub):
    import torch
    from acoustic_isotropic_operators_amd import _capi as A
    print(f"lanes {lanes}, occupancy {occ}, sub {sub}", flush=True)
    c = make_case(seed=40 + sub, nShot=5, sub=sub, dts=0.002 * sub / 2, nts=60)
    g = c.grid
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=c.pad,
                  zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0)
    ctx = A.Context(params)
    ctx.set_velocity(c.vel)
    ctx.set_option("lanes", lanes)
    ctx.set_option("occupancy", occ)
    assert ctx.get_option("lanes") == lanes and ctx.get_option("occupancy") == occ
    srcs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.src_args]
    recs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.rec_args]
    os_, or_ = c.oracle_sources(), c.oracle_receivers()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(lanes * 10 + occ + sub)
    nrec = recs[0].nDev
    # nonlinear
    m = np.zeros((1, 1, c.nts), np.float32)
    m[0, 0] = c.wavelet
    d = torch.zeros((c.nShot, nrec, c.nts), dtype=torch.float32, device=dev)
    ctx.nl_forward(torch.from_numpy(m).to(dev), d, srcs, recs, mem=A.MEM_DEVICE)
    ref = O.nl_forward(g, c.vel, os_, or_, m, "f")
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(d.cpu().numpy(), ref)
    dr = rng.standard_normal((c.nShot, nrec, c.nts)).astype(np.float32)
    ma = torch.zeros((1, 1, c.nts), dtype=torch.float32, device=dev)
    ctx.nl_adjoint(ma, torch.from_numpy(dr).to(dev), srcs, recs, mem=A.MEM_DEVICE)
    np.testing.assert_array_equal(ma.cpu().numpy(), O.nl_adjoint(g, c.vel, os_, or_, dr, 1, "f"))
    # Born
    wav = c.wavelet[None, :].copy()
    refl = rng.standard_normal((c.nx, c.nz)).astype(np.float32)
    db = torch.zeros((c.nShot, nrec, c.nts), dtype=torch.float32, device=dev)
    ctx.born_forward(torch.from_numpy(refl).to(dev), db, srcs, wav, recs, mem=A.MEM_DEVICE)
    np.testing.assert_array_equal(db.cpu().numpy(), O.born_forward(g, c.vel, os_, [wav], or_, refl, "f"))
    img = torch.zeros((c.nx, c.nz), dtype=torch.float32, device=dev)
    ctx.born_adjoint(img, torch.from_numpy(dr).to(dev), srcs, wav, recs, mem=A.MEM_DEVICE)
    np.testing.assert_array_equal(img.cpu().numpy(), O.born_adjoint(g, c.vel, os_, [wav], or_, dr, "f"))

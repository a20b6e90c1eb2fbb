"""aiso_nl_forward_noalloc against aiso_nl_forward through the C-ABI: the
same data, bitwise, on two calls in a row, the second with a smaller input
(fewer shots on a smaller grid); the scratch array keeps the first call's
contents, so the second call shows it clears what it needs.  Also its
refusals: a call whose one shot exceeds the 64 MiB scratch, and host memory.
Collected last, so that a failure here hides no other test under pytest -x."""
import numpy as np
import pytest

from cases import make_case

pytestmark = pytest.mark.gpu


def _setup(A, c):
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=c.pad,
                  zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0)
    ctx = A.Context(params)
    ctx.set_velocity(c.vel)
    srcs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.src_args]
    recs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.rec_args]
    return ctx, srcs, recs


def _both(torch, A, c, seed):
    ctx, srcs, recs = _setup(A, c)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    m = torch.from_numpy(rng.standard_normal((c.nShot, 1, c.nts)).astype(np.float32)).to(dev)
    d0 = torch.from_numpy(rng.standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)).to(dev)
    ref = d0.clone()
    ctx.nl_forward(m, ref, srcs, recs, add=True, mem=A.MEM_DEVICE)
    got = d0.clone()
    ctx.nl_forward_noalloc(m, got, srcs, recs, add=True)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_matches_the_allocating_entry_point_twice():
    import torch
    from acoustic_isotropic_operators_amd import _capi as A
    print("call 1: 5 shots on 58 x 74, sub 2", flush=True)
    ref, got = _both(torch, A, make_case(seed=301, nShot=5, sub=2, dts=0.002, nts=50), 302)
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(got, ref)
    print("call 2: 2 shots on 42 x 58, sub 3", flush=True)
    ref, got = _both(torch, A, make_case(seed=303, nz=42, nx=58, nShot=2, sub=3, dts=0.003, nts=30, pad=6,
                                         shot_dx=10), 304)
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(got, ref)


def test_noalloc_refuses_what_exceeds_its_scratch():
    import torch
    from acoustic_isotropic_operators_amd import _capi as A
    from acoustic_isotropic_operators_amd.workloads import born_lsrtm
    w = born_lsrtm(nShot=1, nts=2000)
    p = w.params
    ctx = A.Context(p)
    ctx.set_velocity(w.vel)
    src = [A.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"]) for iz, ix in w.src]
    # eight receiver rows: ~10,800 regular points x 2000 samples = 86 MB of traces
    nzd, ozd, dzd, nxd, oxd, dxd = w.rec
    rec = [A.Device.grid(8, ozd, 1, nxd, oxd, dxd, p["nz"], p["nx"], p["nts"])]
    print(f"receivers: {rec[0].nDev} devices, {rec[0].nReg} regular points", flush=True)
    dev = torch.device("cuda:0")
    m = torch.zeros((1, 1, p["nts"]), dtype=torch.float32, device=dev)
    d = torch.zeros((1, rec[0].nDev, p["nts"]), dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError, match="64 MiB scratch"):
        ctx.nl_forward_noalloc(m, d, src, rec)

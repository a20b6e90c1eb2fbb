"""The headline workload at its full size on the GPU (Born fwd+adj, 80 shots,
500 x 1200 with a 50-point sponge, nts 2000 at 2 ms, sub 2), through the
C-ABI with device buffers, checked by what does not depend on the size:

* one shot of the full workload, bitwise against the float oracle (the
  oracle takes a few seconds for one shot with OpenMP);
* shots inside the 80-shot call equal the same shots run alone, bitwise;
* linearity in the model, and the dot test <F m, d> = <m, F' d> with FP64
  sums at a stated tolerance;
* a 1000 x 3000 grid (the largest in the serving mix), one shot, with the
  same dot test.
"""
import numpy as np
import pytest

from cases import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    from acoustic_isotropic_operators_amd import _capi as A
    from acoustic_isotropic_operators_amd.workloads import born_lsrtm
    w = born_lsrtm()
    p = w.params
    ctx = A.Context(p)
    ctx.set_velocity(w.vel)
    srcs = [A.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"]) for iz, ix in w.src]
    rec = [A.Device.grid(*w.rec, p["nz"], p["nx"], p["nts"])]
    return torch, A, w, ctx, srcs, rec


def _fwd(torch, A, ctx, w, srcs, rec, m):
    dev = torch.device("cuda:0")
    d = torch.zeros((len(srcs), rec[0].nDev, w.params["nts"]), dtype=torch.float32, device=dev)
    ctx.born_forward(torch.from_numpy(np.ascontiguousarray(m)).to(dev), d, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)
    return d.cpu().numpy()


def _adj(torch, A, ctx, w, srcs, rec, d):
    dev = torch.device("cuda:0")
    img = torch.zeros(w.vel.shape, dtype=torch.float32, device=dev)
    ctx.born_adjoint(img, torch.from_numpy(np.ascontiguousarray(d)).to(dev), srcs, w.wavelet, rec, mem=A.MEM_DEVICE)
    return img.cpu().numpy()


def test_one_shot_of_the_full_workload_matches_the_oracle(setup):
    torch, A, w, ctx, srcs, rec = setup
    p = w.params
    g = O.Grid(nz=p["nz"], nx=p["nx"], nts=p["nts"], sub=p["sub"],
               minPad=min(p["zPadMinus"], p["zPadPlus"], p["xPadMinus"], p["xPadPlus"]), dz=p["dz"], dx=p["dx"],
               dts=p["dts"], alphaCos=p["alphaCos"])
    s = 37
    osrc = [O.Device.grid(1, w.src[s][0], 1, 1, w.src[s][1], 1, p["nz"], p["nx"], p["nts"])]
    orec = [O.Device.grid(*w.rec, p["nz"], p["nx"], p["nts"])]
    print(f"shot {s} of {w.nShot}, grid {p['nz']}x{p['nx']}, nts {p['nts']}", flush=True)
    d = _fwd(torch, A, ctx, w, srcs[s:s + 1], rec, w.model)
    ref = O.born_forward(g, w.vel, osrc, [w.wavelet], orec, w.model, "f")
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(d, ref)
    img = _adj(torch, A, ctx, w, srcs[s:s + 1], rec, d)
    np.testing.assert_array_equal(img, O.born_adjoint(g, w.vel, osrc, [w.wavelet], orec, d, "f"))


def test_shots_in_the_full_call_equal_shots_alone(setup):
    torch, A, w, ctx, srcs, rec = setup
    d = _fwd(torch, A, ctx, w, srcs, rec, w.model)
    for s in (0, 41, 79):
        np.testing.assert_array_equal(d[s:s + 1], _fwd(torch, A, ctx, w, srcs[s:s + 1], rec, w.model))
    img_all = _adj(torch, A, ctx, w, srcs, rec, d)
    img_sum = np.zeros_like(img_all, dtype=np.float64)
    for s in range(0, w.nShot, 10):         # the image is the sum over shots of the per-shot images
        img_sum += _adj(torch, A, ctx, w, srcs[s:s + 10], rec, d[s:s + 10])
    assert rel_l2(img_all, img_sum) < 1e-6


def test_linearity_and_dot_product_at_full_size(setup):
    torch, A, w, ctx, srcs, rec = setup
    rng = np.random.default_rng(7)
    m1 = w.model
    m2 = (rng.standard_normal(w.vel.shape) * (w.model != 0)).astype(np.float32)
    d1 = _fwd(torch, A, ctx, w, srcs, rec, m1)
    d2 = _fwd(torch, A, ctx, w, srcs, rec, m2)
    d12 = _fwd(torch, A, ctx, w, srcs, rec, (2.0 * m1 - 0.5 * m2).astype(np.float32))
    lin = rel_l2(d12, 2.0 * d1.astype(np.float64) - 0.5 * d2.astype(np.float64))
    dr = rng.standard_normal(d1.shape).astype(np.float32)
    a = float(np.vdot(d1.astype(np.float64).ravel(), dr.astype(np.float64).ravel()))
    b = float(np.vdot(m1.astype(np.float64).ravel(), _adj(torch, A, ctx, w, srcs, rec, dr).astype(np.float64).ravel()))
    print(f"linearity relL2 {lin:.2e}; <Fm,d> {a:.9e} <m,F'd> {b:.9e} rel {abs(a - b) / abs(b):.2e}")
    # float rounding over 2 x 3998 fine steps: 1.2e-5 (linearity) and 8.5e-6
    # (dot test) measured on MI355X; a model-dependent term would show at
    # order 1.  The dot test holds to the 1e-5 bar of the operator
    # (F/include/operator.h:50-88).
    assert lin < 1e-4
    assert abs(a - b) <= 1e-5 * abs(b)


def test_largest_grid_one_shot_dot_product():
    import torch
    from acoustic_isotropic_operators_amd import _capi as A
    from acoustic_isotropic_operators_amd.workloads import pad_model, ricker
    nz, nx, pad, nts, dts, sub = 1000, 3000, 50, 400, 0.002, 2
    rng = np.random.default_rng(11)
    v = np.broadcast_to(1500.0 + 2500.0 * np.arange(nz) / (nz - 1), (nx, nz)).astype(np.float32)
    vp, zPlus, xPlus = pad_model(v, pad, pad)
    NX, NZ = vp.shape
    params = dict(nz=NZ, nx=NX, dz=10.0, dx=10.0, nts=nts, dts=dts, sub=sub, zPadMinus=pad, zPadPlus=zPlus,
                  xPadMinus=pad, xPadPlus=xPlus, alphaCos=0.95, fMax=37.5)
    ctx = A.Context(params)
    ctx.set_velocity(vp)
    iz = 7 + pad
    srcs = [A.Device.grid(1, iz, 1, 1, 5 + pad + nx // 2, 1, NZ, NX, nts)]
    rec = [A.Device.grid(1, iz, 1, nx, 5 + pad, 1, NZ, NX, nts)]
    wav = ricker(nts, dts, 15.0, 0.1)[None, :]
    m = rng.standard_normal((NX, NZ)).astype(np.float32)
    dev = torch.device("cuda:0")
    d = torch.zeros((1, nx, nts), dtype=torch.float32, device=dev)
    ctx.born_forward(torch.from_numpy(m).to(dev), d, srcs, wav, rec, mem=A.MEM_DEVICE)
    dr = torch.from_numpy(rng.standard_normal((1, nx, nts)).astype(np.float32)).to(dev)
    img = torch.zeros((NX, NZ), dtype=torch.float32, device=dev)
    ctx.born_adjoint(img, dr, srcs, wav, rec, mem=A.MEM_DEVICE)
    a = float(np.vdot(d.cpu().numpy().astype(np.float64).ravel(), dr.cpu().numpy().astype(np.float64).ravel()))
    b = float(np.vdot(m.astype(np.float64).ravel(), img.cpu().numpy().astype(np.float64).ravel()))
    print(f"{NZ}x{NX}: <Fm,d> {a:.9e} <m,F'd> {b:.9e} rel {abs(a - b) / abs(b):.2e}")
    assert np.isfinite(a) and abs(a) > 0
    assert abs(a - b) <= 1e-5 * abs(b)

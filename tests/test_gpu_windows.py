"""The batched form over windows of uneven size (aiso_batch_*) against plain
one-shot launches and the float oracle.

Each window's data must equal, bitwise, what aiso_nl_forward gives for that
shot alone on a context of that window (one shot per launch), for every tile
shape, windows-per-launch split, persistent or one-tile-per-workgroup launch,
with and without graph capture, on a first call and on graph replays; two
windows are also checked against the float oracle.  At full size (128
windows of width 300-1500 at nz 400, nts 3000) every window is compared with
its plain launch after the autotuner has picked the configuration.

Each test prints what it runs before it runs it.
"""
import numpy as np
import pytest

from cases import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _grid(p):
    return O.Grid(nz=p["nz"], nx=p["nx"], nts=p["nts"], sub=p["sub"],
                  minPad=min(p["zPadMinus"], p["zPadPlus"], p["xPadMinus"], p["xPadPlus"]), dz=p["dz"], dx=p["dx"],
                  dts=p["dts"], alphaCos=p["alphaCos"])


def _devices(A, w, k):
    NX, NZ = w.vels[k].shape
    nts = w.params["nts"]
    iz, ix = w.src[k]
    src = A.Device.grid(1, iz, 1, 1, ix, 1, NZ, NX, nts)
    rec = A.Device.grid(*w.rec[k], NZ, NX, nts)
    return src, rec


def _plain(A, w, k, model):
    """The window's shot alone through aiso_nl_forward, one shot per launch."""
    p = w.window_params(k)
    ctx = A.Context(p)
    ctx.set_batch(1)
    ctx.set_velocity(w.vels[k])
    src, rec = _devices(A, w, k)
    d = np.zeros((1, rec.nDev, p["nts"]), np.float32)
    ctx.nl_forward(model[None], d, [src], [rec])
    return d[0]


def _batch(A, w):
    b = A.Batch(w.params)
    devs = [_devices(A, w, k) for k in range(w.nShot)]
    b.set_windows(w.vels, [s for s, _ in devs], [r for _, r in devs])
    return b, devs


@pytest.fixture(scope="module")
def small():
    from acoustic_isotropic_operators_amd import _capi as A
    from acoustic_isotropic_operators_amd.workloads import batched_windows
    w = batched_windows(nShot=7, seed=11, nts=160, dts=0.002, sub=2, nx_range=(60, 230), nz=70)
    model = w.wavelet.astype(np.float32)                         # [1][nts], one source point
    print(f"small windows: {[v.shape for v in w.vels]}", flush=True)
    ref = [_plain(A, w, k, model) for k in range(w.nShot)]
    return A, w, model, ref


@pytest.mark.parametrize("tz,tx,spl,persist,graph,lanes", [
    (32, 16, 0, 0, True, 1), (32, 32, 3, 1, True, 2), (64, 16, 2, 0, False, 2), (64, 32, 0, 2, True, 1),
    (32, 16, 4, 1, False, 1), (32, 32, 1, 0, True, 3), (32, 16, 2, 0, True, 2),
])
def test_windows_match_plain_launches(small, tz, tx, spl, persist, graph, lanes):
    A, w, model, ref = small
    print(f"tile {tz}x{tx}, windows per launch {spl}, workgroups per CU {persist}, graph {graph}, lanes {lanes}",
          flush=True)
    b, _devs = _batch(A, w)
    b.set_config(tz, tx, spl, persist, graph, lanes)
    # a graph is captured on the first call and replayed by the next three,
    # with eager work of the program (a plain call and a device sync) between
    for call in range(4):
        out = [np.full((r.shape[0], r.shape[1]), np.nan, np.float32) for r in ref]
        b.nl_forward([model], out)
        for k in range(w.nShot):
            assert np.abs(ref[k]).max() > 0
            np.testing.assert_array_equal(out[k], ref[k], err_msg=f"window {k}, call {call}")
        np.testing.assert_array_equal(_plain(A, w, call % w.nShot, model), ref[call % w.nShot])
    info = b.info()
    print(f"info {info}", flush=True)
    assert info["windows"] == w.nShot


def test_windows_match_the_oracle(small):
    A, w, model, ref = small
    for k in (0, w.nShot - 1):
        p = w.window_params(k)
        print(f"window {k}: {p['nz']}x{p['nx']} against the float oracle", flush=True)
        iz, ix = w.src[k]
        osrc = [O.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"])]
        orec = [O.Device.grid(*w.rec[k], p["nz"], p["nx"], p["nts"])]
        of = O.nl_forward(_grid(p), w.vels[k], osrc, orec, model[None], "f")[0]
        od = O.nl_forward(_grid(p), w.vels[k], osrc, orec, model[None].astype(np.float64), "d")[0]
        print(f"relL2 vs float oracle {rel_l2(ref[k], of):.3e}, vs FP64 {rel_l2(ref[k], od):.3e}", flush=True)
        np.testing.assert_array_equal(ref[k], of)


def test_windows_device_memory_add_and_per_window_models(small):
    import torch
    A, w, model, ref = small
    print("device operands, add, one model per window", flush=True)
    b, _devs = _batch(A, w)
    dev = torch.device("cuda:0")
    scale = [np.float32(1 + k) for k in range(w.nShot)]
    models = [torch.from_numpy(model * s).to(dev) for s in scale]
    base = [torch.ones(r.shape, dtype=torch.float32, device=dev) for r in ref]
    b.nl_forward(models, base, add=True, mem=A.MEM_DEVICE)
    for k in range(w.nShot):
        want = _plain(A, w, k, model * scale[k]) + np.float32(1)
        np.testing.assert_array_equal(base[k].cpu().numpy(), want, err_msg=f"window {k}")


def test_windows_refuse_a_device_in_the_fat_ring():
    from acoustic_isotropic_operators_amd import _capi as A
    from acoustic_isotropic_operators_amd.workloads import batched_windows
    w = batched_windows(nShot=2, seed=3, nts=40, nx_range=(60, 80), nz=40)
    b = A.Batch(w.params)
    NX, NZ = w.vels[0].shape
    devs = [_devices(A, w, k) for k in range(2)]
    ring = A.Device.grid(1, 2, 1, 1, 40, 1, NZ, NX, w.params["nts"])          # iz = 2 < FAT
    b.set_windows(w.vels, [ring, devs[1][0]], [d[1] for d in devs])
    out = [np.zeros((d[1].nDev, w.params["nts"]), np.float32) for d in devs]
    with pytest.raises(RuntimeError, match="FAT ring"):
        b.nl_forward([w.wavelet.astype(np.float32)], out)


def test_windows_full_size_after_tuning():
    """128 windows (nz 400, widths 300-1500), nts 3000 at 2 ms: the tuned,
    graph-captured batch against the plain one-shot launches, every window."""
    import time
    import torch
    from acoustic_isotropic_operators_amd import _capi as A
    from acoustic_isotropic_operators_amd.workloads import batched_windows
    w = batched_windows()
    print(f"{w.name}: {w.nShot} windows", flush=True)
    dev = torch.device("cuda:0")
    model = torch.from_numpy(w.wavelet.astype(np.float32)).to(dev)
    b, _devs = _batch(A, w)
    t0 = time.time()
    cfg = b.tune(steps=12)
    print(f"tuned in {time.time() - t0:.2f} s: {cfg}", flush=True)
    out = [torch.zeros((d[1].nDev, w.params["nts"]), dtype=torch.float32, device=dev) for d in _devs]
    b.nl_forward([model], out, mem=A.MEM_DEVICE)
    print(f"batched call: {b.info()}", flush=True)
    mismatch = []
    for k in range(w.nShot):
        p = w.window_params(k)
        ctx = A.Context(p)
        ctx.set_batch(1)
        ctx.set_velocity(w.vels[k])
        src, rec = _devices(A, w, k)
        d = torch.zeros((1, rec.nDev, p["nts"]), dtype=torch.float32, device=dev)
        ctx.nl_forward(model[None], d, [src], [rec], mem=A.MEM_DEVICE)
        a, e = out[k].cpu().numpy(), d[0].cpu().numpy()
        if not np.array_equal(a, e):
            mismatch.append((k, rel_l2(a, e)))
        if k % 32 == 0:
            print(f"window {k} ({p['nx']} wide) compared", flush=True)
    print(f"windows differing from their plain launch: {mismatch}", flush=True)
    assert not mismatch

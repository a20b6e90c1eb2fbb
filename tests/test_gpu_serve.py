"""The request server on the GPU: requests of several callers batched into
shared engine calls give, each, what the same shot gives alone (bitwise),
for every request kind, with two resident models and interleaved arrivals,
both with batches run one at a time and with concurrent slots (continuous
batching: calls overlapping on the device while later requests arrive);
the engine's per-shot image option returns the images whose sum the plain
adjoint returns."""
import numpy as np
import pytest

from cases import make_case
from acoustic_isotropic_operators_amd.serve import Request, Server

pytestmark = pytest.mark.gpu


def _models():
    out = []
    for k, (seed, nExt) in enumerate([(201, 5), (202, 5)]):
        c = make_case(seed=seed, nShot=1, nts=40, sub=2, dts=0.002)
        params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=c.pad,
                      zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0)
        out.append((f"m{k}", c, params))
    return out


def _requests(srv, rng, n=24):
    cases = {mid: (c, params) for mid, c, params in _models()}
    refl = {mid: rng.standard_normal((c.nx, c.nz)).astype(np.float32) for mid, (c, _) in cases.items()}
    kinds = ["nl_forward", "born_forward", "born_adjoint", "bornext_adjoint"]
    reqs = []
    for k in range(n):
        mid = f"m{k % 2}"
        c, params = cases[mid]
        kind = kinds[(k // 2) % 4]
        r = Request(kind, mid, (16, 15 + 3 * (k % 7)))
        if kind == "born_forward":
            r.payload, r.reflectivity_id = refl[mid], mid
        elif kind != "nl_forward":
            r.payload = rng.standard_normal((srv.models[mid].rec[0].nDev, c.nts)).astype(np.float32)
        reqs.append(r)
    return reqs


def _server(**kw):
    srv = Server(max_models=2, nExt=5, **kw)
    for mid, c, params in _models():
        srv.add_model(mid, params, c.vel, c.rec_args[0][1], c.wavelet)
    return srv


def _check_against_alone(reqs):
    """Each request against the same shot alone, on a fresh context."""
    from acoustic_isotropic_operators_amd import _capi as A
    for mid, c, params in _models():
        ctx = A.Context(dict(params, nExt=5, extension=A.EXT_OFFSET))
        ctx.set_velocity(c.vel)
        rec = [A.Device.grid(*c.rec_args[0][1], c.nz, c.nx, c.nts)]
        wav = c.wavelet[None, :].astype(np.float32)
        for r in (r for r in reqs if r.model == mid):
            assert r.error is None, r.error
            src = [A.Device.grid(1, r.source[0], 1, 1, r.source[1], 1, c.nz, c.nx, c.nts)]
            if r.kind == "nl_forward":
                d = np.zeros((1, rec[0].nDev, c.nts), np.float32)
                ctx.nl_forward(wav[None], d, src, rec)
                ref = d[0]
            elif r.kind == "born_forward":
                d = np.zeros((1, rec[0].nDev, c.nts), np.float32)
                ctx.born_forward(r.payload, d, src, wav, rec)
                ref = d[0]
            elif r.kind == "born_adjoint":
                ref = np.zeros((c.nx, c.nz), np.float32)
                ctx.born_adjoint(ref, r.payload[None], src, wav, rec)
            else:
                ref = np.zeros((5, c.nx, c.nz), np.float32)
                ctx.bornext_adjoint(ref, r.payload[None], src, wav, rec)
            assert np.abs(ref).max() > 0
            np.testing.assert_array_equal(r.result, ref, err_msg=f"{mid} {r.kind} {r.source}")


def test_batched_requests_equal_single_shots():
    print("synchronous batches, 24 requests on 2 models", flush=True)
    srv = _server(max_shots=4)
    reqs = _requests(srv, np.random.default_rng(203))
    for r in reqs:
        srv.submit(r)
    while srv.pending():
        srv.step()
    assert srv.calls < len(reqs)                     # requests did share calls
    _check_against_alone(reqs)


def test_continuous_batching_results_equal_single_shots():
    """Three slots: calls run concurrently on their own contexts and streams
    while later requests arrive; every result equals its shot alone."""
    import time
    print("continuous batching, 3 slots, 40 requests in 4 waves", flush=True)
    srv = _server(max_shots=3, slots=3)
    reqs = _requests(srv, np.random.default_rng(206), n=40)
    srv.start()
    try:
        for w in range(4):
            for r in reqs[w * 10:(w + 1) * 10]:
                srv.submit(r)
            time.sleep(0.01)                         # the next wave arrives while calls run
        srv.drain()
    finally:
        srv.stop()
    print(f"calls {srv.calls}, most calls running at once {srv.max_running}", flush=True)
    assert srv.max_running >= 2
    assert srv.calls < len(reqs)
    _check_against_alone(reqs)


def test_image_per_shot_option():
    from acoustic_isotropic_operators_amd import _capi as A
    c = make_case(seed=204, nShot=3, nts=40, sub=2, dts=0.002, nExt=3)
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=c.pad,
                  zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0, nExt=3,
                  extension=A.EXT_OFFSET)
    ctx = A.Context(params)
    ctx.set_velocity(c.vel)
    srcs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.src_args]
    rec = [A.Device.grid(*c.rec_args[0][1], c.nz, c.nx, c.nts)]
    wav = c.wavelet[None, :].astype(np.float32)
    d = np.random.default_rng(205).standard_normal((3, rec[0].nDev, c.nts)).astype(np.float32)
    for ext in (False, True):
        shape = (3, c.nx, c.nz) if ext else (c.nx, c.nz)
        fn = ctx.bornext_adjoint if ext else ctx.born_adjoint
        summed = np.zeros(shape, np.float32)
        fn(summed, d, srcs, wav, rec)
        ctx.set_option("imagePerShot", 1)
        per = np.zeros((3,) + shape, np.float32)
        fn(per, d, srcs, wav, rec)
        ctx.set_option("imagePerShot", 0)
        for s in range(3):
            one = np.zeros(shape, np.float32)
            fn(one, d[s:s + 1], srcs[s:s + 1], wav, rec)
            np.testing.assert_array_equal(per[s], one)
        np.testing.assert_array_equal(((per[0] + per[1]) + per[2]), summed)

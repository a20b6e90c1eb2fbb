"""The request server at the serving configuration's sizes: the four resident
models of the serving mix (200 x 400 at nts 4000 up to 1000 x 3000 at nts
1000, receivers on every column) and every request kind (nonlinear data,
Born data, Born image, 21-offset extended image), two requests per (model,
kind), all queued at once and served with continuous batching over four
slots.  Each result is bitwise what its shot gives alone in a plain one-shot
call; the smallest model's nonlinear and Born data are also bitwise the
float oracle's."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KINDS = ["nl_forward", "born_forward", "born_adjoint", "bornext_adjoint"]


def _grid(p):
    return O.Grid(nz=p["nz"], nx=p["nx"], nts=p["nts"], sub=p["sub"],
                  minPad=min(p["zPadMinus"], p["zPadPlus"], p["xPadMinus"], p["xPadPlus"]), dz=p["dz"], dx=p["dx"],
                  dts=p["dts"], alphaCos=p["alphaCos"])


def test_every_kind_on_every_serving_grid_equals_its_shot_alone():
    from acoustic_isotropic_operators_amd import _capi as A
    from acoustic_isotropic_operators_amd.serve import Request, Server
    from acoustic_isotropic_operators_amd.workloads import serving_models
    models = serving_models()
    srv = Server(max_models=4, max_shots=32, slots=4, nExt=21)
    rng = np.random.default_rng(55)
    refl, reqs = {}, []
    for mid, params, vel, rec, wav, (iz, x0, x1) in models:
        print(f"model {mid}: padded {params['nz']}x{params['nx']}, nts {params['nts']}", flush=True)
        srv.add_model(mid, params, vel, rec, wav)
        r_ = np.zeros((params["nx"], params["nz"]), np.float32)
        r_[x0:x1, iz + 40:iz + 43] = 0.1
        refl[mid] = r_
        for kind in KINDS:
            for ix in (x0 + (x1 - x0) // 3, x0 + 2 * (x1 - x0) // 3):
                r = Request(kind, mid, (iz, ix))
                if kind == "born_forward":
                    r.payload, r.reflectivity_id = refl[mid], mid
                elif kind != "nl_forward":
                    r.payload = rng.standard_normal((srv.models[mid].rec[0].nDev, params["nts"])).astype(np.float32)
                reqs.append(r)
    print(f"serving {len(reqs)} requests", flush=True)
    srv.start()
    try:
        for r in reqs:
            srv.submit(r)
        srv.drain()
    finally:
        srv.stop()
    print(f"calls {srv.calls}, most calls running at once {srv.max_running}", flush=True)
    assert srv.calls < len(reqs)
    for mid, params, vel, rec, wav, _ in models:
        p = dict(params, nExt=21, extension=A.EXT_OFFSET)
        ctx = A.Context(p)
        ctx.set_velocity(vel)
        recs = [A.Device.grid(*rec, p["nz"], p["nx"], p["nts"])]
        nrec, nts = recs[0].nDev, p["nts"]
        w1 = np.ascontiguousarray(wav, np.float32).reshape(1, -1)
        for r in (r for r in reqs if r.model == mid):
            print(f"{mid} {r.kind} at {r.source}: against the shot alone", flush=True)
            assert r.error is None, r.error
            src = [A.Device.grid(1, r.source[0], 1, 1, r.source[1], 1, p["nz"], p["nx"], nts)]
            if r.kind == "nl_forward":
                d = np.zeros((1, nrec, nts), np.float32)
                ctx.nl_forward(w1[None], d, src, recs)
                ref = d[0]
            elif r.kind == "born_forward":
                d = np.zeros((1, nrec, nts), np.float32)
                ctx.born_forward(refl[mid], d, src, w1, recs)
                ref = d[0]
            elif r.kind == "born_adjoint":
                ref = np.zeros((p["nx"], p["nz"]), np.float32)
                ctx.born_adjoint(ref, r.payload[None], src, w1, recs)
            else:
                ref = np.zeros((21, p["nx"], p["nz"]), np.float32)
                ctx.bornext_adjoint(ref, r.payload[None], src, w1, recs)
            assert np.abs(ref).max() > 0
            np.testing.assert_array_equal(r.result, ref, err_msg=f"{mid} {r.kind} {r.source}")
        ctx.close()
    # the smallest grid's data against the float oracle
    mid, params, vel, rec, wav, _ = models[0]
    g = _grid(params)
    orec = [O.Device.grid(*rec, params["nz"], params["nx"], params["nts"])]
    w1 = np.ascontiguousarray(wav, np.float32).reshape(1, -1)
    first_of_kind = [r for r in reqs if r.model == mid and r.kind in ("nl_forward", "born_forward")][::2]
    for r in first_of_kind:
        print(f"{mid} {r.kind} at {r.source}: against the float oracle", flush=True)
        osrc = [O.Device.grid(1, r.source[0], 1, 1, r.source[1], 1, params["nz"], params["nx"], params["nts"])]
        if r.kind == "nl_forward":
            of = O.nl_forward(g, vel, osrc, orec, w1[None], "f")[0]
        else:
            of = O.born_forward(g, vel, osrc, [w1], orec, refl[mid], "f")[0]
        np.testing.assert_array_equal(r.result, of)

"""The request server's scheduling on the CPU, with the engine replaced by a
recorder: batches form by (model, kind, reflectivity) in arrival order of
their oldest request, respect the shot cap and the HBM budget, leave other
requests queued in order; Born forward requests without a reflectivity id
share a call only with the same array; eviction spares models with pending
requests; with worker threads, a request that arrives while a call runs is
admitted into a free slot and runs concurrently."""
import threading
import time

import numpy as np
import pytest

from acoustic_isotropic_operators_amd import serve


class _FakeModel:
    closed = []

    def __init__(self, params, vel, receivers, wavelet, nExt, device, options=None):
        self.params = params
        self.options = options
        self.bytes_per_shot = {k: params.get("bytes", 1) for k in serve.KINDS}
        self.runs = []
        self.active = 0
        self.slots = []

    def new_context(self):
        return object()

    def close(self):
        _FakeModel.closed.append(self.params["name"])


@pytest.fixture
def server(monkeypatch):
    monkeypatch.setattr(serve, "_Model", _FakeModel)
    monkeypatch.setattr(serve.Server, "_run", lambda self, m, ctx, batch: m.runs.append([id(r) for r in batch]))
    _FakeModel.closed = []
    return serve.Server(max_models=4, hbm_budget_gb=4e-9, max_shots=3)


def test_batches_follow_the_oldest_request(server):
    server.budget = 1.0               # one shot of one byte at a time
    for name in ("a", "b"):
        server.add_model(name, {"name": name, "bytes": 1}, None, None, None)
    R = serve.Request
    z = np.zeros(1)
    reqs = [R("born_adjoint", "a", (1, 1)), R("nl_forward", "b", (1, 2)), R("born_adjoint", "a", (1, 3)),
            R("born_forward", "a", (1, 4), payload=z, reflectivity_id=1),
            R("born_forward", "a", (1, 5), payload=z, reflectivity_id=2),
            R("born_adjoint", "a", (1, 6)), R("born_adjoint", "a", (1, 7)), R("nl_forward", "b", (1, 8))]
    for r in reqs:
        server.submit(r)
    done = [server.step() for _ in range(6)]
    assert all(len(b) == 1 for b in done)
    assert [b[0] for b in done] == reqs[:6]
    assert server.pending() == 2


def test_shot_cap_and_keys(server):
    server.budget = 1e9
    server.add_model("a", {"name": "a", "bytes": 1}, None, None, None)
    R = serve.Request
    z = np.zeros(1)
    reqs = [R("born_adjoint", "a", (1, k)) for k in range(5)] + [R("born_forward", "a", (1, 9), payload=z,
                                                                   reflectivity_id=7)]
    reqs.insert(2, R("born_forward", "a", (1, 8), payload=z, reflectivity_id=7))
    for r in reqs:
        server.submit(r)
    b1 = server.step()
    assert [r.source[1] for r in b1] == [0, 1, 2]            # cap 3, same key only, in order
    b2 = server.step()
    assert [r.source[1] for r in b2] == [8, 9]               # the reflectivity's two requests together
    b3 = server.step()
    assert [r.source[1] for r in b3] == [3, 4]
    assert server.pending() == 0
    with pytest.raises(KeyError):
        server.submit(R("nl_forward", "zzz", (1, 1)))
    with pytest.raises(ValueError):
        server.submit(R("tomo", "a", (1, 1)))
    with pytest.raises(ValueError):
        server.submit(R("born_forward", "a", (1, 1)))        # no reflectivity


def test_born_forward_without_an_id_shares_calls_only_with_the_same_array(server):
    server.budget = 1e9
    server.add_model("a", {"name": "a", "bytes": 1}, None, None, None)
    R = serve.Request
    m1, m2 = np.zeros(4), np.ones(4)
    reqs = [R("born_forward", "a", (1, 0), payload=m1), R("born_forward", "a", (1, 1), payload=m2),
            R("born_forward", "a", (1, 2), payload=m1)]
    for r in reqs:
        server.submit(r)
    b1, b2 = server.step(), server.step()
    assert [r.source[1] for r in b1] == [0, 2]
    assert [r.source[1] for r in b2] == [1]


def test_budget_holds_back_what_does_not_fit(server):
    server.budget = 5.0
    server.add_model("a", {"name": "a", "bytes": 2}, None, None, None)
    R = serve.Request
    for k in range(4):
        server.submit(R("nl_forward", "a", (1, k)))
    server.slots = 2
    with server.cv:
        first = server._take()
        assert len(first[2]) == 2 and server.resident_bytes == 4.0
        assert server._take() is None                        # 1 byte left, a shot needs 2, a call is running
    server._execute(first)
    assert server.resident_bytes == 4.0                      # the context keeps its buffers
    assert len(server.step()) == 2                           # ...and reuses them
    assert server.resident_bytes == 4.0 and len(server.models["a"].slots) == 1
    server.budget = 1.0                                      # alone, a batch beyond the budget still runs
    server.submit(serve.Request("nl_forward", "a", (1, 9)))
    assert len(server.step()) == 1


def test_least_recently_used_idle_model_is_evicted(server):
    for name in ("a", "b", "c", "d"):
        server.add_model(name, {"name": name}, None, None, None)
    server.budget = 1e9
    server.submit(serve.Request("nl_forward", "a", (1, 1)))
    server.step()                                            # "a" becomes the most recent
    server.submit(serve.Request("nl_forward", "b", (1, 1)))  # "b" has a pending request
    server.add_model("e", {"name": "e"}, None, None, None)
    assert _FakeModel.closed == ["c"]
    assert list(server.models) == ["b", "d", "a", "e"]
    for name in ("d", "a", "e"):
        server.submit(serve.Request("nl_forward", name, (1, 1)))
    with pytest.raises(RuntimeError, match="evicted"):
        server.add_model("f", {"name": "f"}, None, None, None)
    assert server.pending() == 4


def test_worker_slots_overlap_calls(monkeypatch):
    """Two slots: a request that arrives while the first call runs starts at
    once in the second slot; the first call finishing is not awaited."""
    monkeypatch.setattr(serve, "_Model", _FakeModel)
    gate = threading.Event()
    started = []

    def run(self, m, ctx, batch):
        started.append([r.source[1] for r in batch])
        if batch[0].source[1] == 0:
            assert gate.wait(5.0)                            # the first call runs until released
        for r in batch:
            r.result = np.full(2, r.source[1], np.float32)

    monkeypatch.setattr(serve.Server, "_run", run)
    srv = serve.Server(max_models=2, hbm_budget_gb=1.0, max_shots=4, slots=2)
    srv.add_model("a", {"name": "a", "bytes": 1}, None, None, None)
    srv.start()
    try:
        srv.submit(serve.Request("nl_forward", "a", (1, 0)))
        t0 = time.time()
        while not started and time.time() - t0 < 5:
            time.sleep(0.001)
        late = serve.Request("born_adjoint", "a", (1, 1), payload=np.zeros(1))
        srv.submit(late)
        t0 = time.time()
        while late.done is None and time.time() - t0 < 5:
            time.sleep(0.001)
        assert late.done is not None and late.error is None   # done while the first call still runs
        assert srv.running == 1 and srv.max_running == 2
        gate.set()
        srv.drain()
    finally:
        gate.set()
        srv.stop()
    assert started == [[0], [1]] and srv.calls == 2


def test_errors_stay_with_their_requests(monkeypatch):
    monkeypatch.setattr(serve, "_Model", _FakeModel)

    def run(self, m, ctx, batch):
        if batch[0].kind == "born_adjoint":
            raise RuntimeError("engine refused")
        for r in batch:
            r.result = np.zeros(1)

    monkeypatch.setattr(serve.Server, "_run", run)
    srv = serve.Server(max_models=1, hbm_budget_gb=1.0, slots=2)
    srv.add_model("a", {"name": "a"}, None, None, None)
    bad = serve.Request("born_adjoint", "a", (1, 1), payload=np.zeros(1))
    good = serve.Request("nl_forward", "a", (1, 2))
    srv.start()
    try:
        srv.submit(bad)
        srv.submit(good)
        srv.drain()
    finally:
        srv.stop()
    assert "engine refused" in bad.error and good.error is None and good.result is not None


def test_engine_options_reach_every_model(monkeypatch):
    monkeypatch.setattr(serve, "_Model", _FakeModel)
    srv = serve.Server(max_models=2, options={"lanes": 1, "occupancy": 8})
    srv.add_model("a", {"name": "a"}, None, None, None)
    assert srv.models["a"].options == {"lanes": 1, "occupancy": 8}


def test_stream_latency_accounting(server):
    server.add_model("a", {"name": "a"}, None, None, None)
    server.budget = 1e9
    t = [0.0]

    def clock():
        t[0] += 0.25
        return t[0]

    reqs = [serve.Request("nl_forward", "a", (1, k), arrival=0.1 * k) for k in range(4)]
    wall = serve.run_stream(server, reqs, clock=clock)
    # arrivals and completions on the same (fake) clock: every tick is 0.25
    assert all(r.done is not None and r.done >= r.arrival for r in reqs)
    assert max(r.done for r in reqs) < 10.0
    assert wall == pytest.approx(max(r.done for r in reqs) - 0.25)
    assert np.isfinite(wall)

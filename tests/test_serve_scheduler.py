"""The request server's scheduling on the CPU, with the engine replaced by a
recorder: batches form by (model, kind, reflectivity) in arrival order of
their oldest request, respect the shot cap, leave other requests queued in
order, and loading a fifth model evicts the least recently used one."""
import numpy as np
import pytest

from acoustic_isotropic_operators_amd import serve


class _FakeModel:
    closed = []

    def __init__(self, params, vel, receivers, wavelet, nExt, device):
        self.params = params
        self.bytes_per_shot = {k: params.get("bytes", 1) for k in serve.KINDS}
        self.runs = []
        self.ctx = self

    def close(self):
        _FakeModel.closed.append(self.params["name"])


@pytest.fixture
def server(monkeypatch):
    monkeypatch.setattr(serve, "_Model", _FakeModel)
    monkeypatch.setattr(serve.Server, "_run", lambda self, m, batch: m.runs.append([id(r) for r in batch]))
    _FakeModel.closed = []
    return serve.Server(max_models=4, hbm_budget_gb=4e-9, max_shots=3)


def test_batches_follow_the_oldest_request(server):
    for name in ("a", "b"):
        server.add_model(name, {"name": name, "bytes": 1}, None, None, None)
    R = serve.Request
    reqs = [R("born_adjoint", "a", (1, 1)), R("nl_forward", "b", (1, 2)), R("born_adjoint", "a", (1, 3)),
            R("born_forward", "a", (1, 4), reflectivity_id=1), R("born_forward", "a", (1, 5), reflectivity_id=2),
            R("born_adjoint", "a", (1, 6)), R("born_adjoint", "a", (1, 7)), R("nl_forward", "b", (1, 8))]
    for r in reqs:
        server.submit(r)
    done = [server.step() for _ in range(6)]
    # budget 4 bytes shared by 4 models: one shot of 1 byte per model share -> cap 1
    assert all(len(b) == 1 for b in done)
    assert [b[0] for b in done] == reqs[:6]
    assert server.pending() == 2


def test_shot_cap_and_keys(server):
    server.budget = 1e9
    server.add_model("a", {"name": "a", "bytes": 1}, None, None, None)
    R = serve.Request
    reqs = [R("born_adjoint", "a", (1, k)) for k in range(5)] + [R("born_forward", "a", (1, 9), reflectivity_id=7)]
    reqs.insert(2, R("born_forward", "a", (1, 8), reflectivity_id=7))
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


def test_least_recently_used_model_is_evicted(server):
    for name in ("a", "b", "c", "d"):
        server.add_model(name, {"name": name}, None, None, None)
    server.submit(serve.Request("nl_forward", "a", (1, 1)))
    server.step()                                            # "a" becomes the most recent
    server.add_model("e", {"name": "e"}, None, None, None)
    assert _FakeModel.closed == ["b"]
    assert list(server.models) == ["c", "d", "a", "e"]


def test_stream_latency_accounting(server):
    server.add_model("a", {"name": "a"}, None, None, None)
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

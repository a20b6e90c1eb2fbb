"""Request server over the engine: independent one-shot modeling requests
from many callers, batched into multi-shot engine calls.

Request kinds, each one shot with a point source on a resident velocity
model's grid and that model's receiver line:

* ``nl_forward``: nonlinear data of the model's wavelet;
* ``born_forward``: Born data of a reflectivity (requests that name the same
  ``reflectivity_id`` share one call);
* ``born_adjoint``: the shot's Born image of its data;
* ``bornext_adjoint``: the shot's subsurface-offset extended image
  (``nExt`` offsets, 21 by default) of its data.

Scheduling: first come, first served by batch.  The oldest pending request
picks the batch key (model, kind, reflectivity); every pending request with
that key joins it, up to ``max_shots`` and up to what the HBM budget admits
for that model (an estimate of the engine's bytes per shot times the shots,
within an equal share of the budget per resident model).  The adjoint
kinds run with the engine's ``imagePerShot`` option, so each request gets its
own image from the shared call.  At most ``max_models`` velocity models are
resident; loading another evicts the least recently used one.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import _capi as A

KINDS = ("nl_forward", "born_forward", "born_adjoint", "bornext_adjoint")


@dataclasses.dataclass
class Request:
    kind: str
    model: object                        # resident velocity model id
    source: tuple                        # (iz, ix) on the model's padded grid
    payload: np.ndarray | None = None    # born_forward: reflectivity [nx][nz]; adjoints: data [nRec][nts]
    reflectivity_id: object = None
    arrival: float = 0.0
    done: float | None = None
    result: np.ndarray | None = None


class _Model:
    def __init__(self, params, vel, receivers, wavelet, nExt, device):
        p = dict(params, nExt=nExt, extension=A.EXT_OFFSET)
        self.params = p
        self.ctx = A.Context(p, device)
        self.ctx.set_velocity(np.ascontiguousarray(vel, np.float32))
        self.ctx.set_option("imagePerShot", 1)
        self.rec = [A.Device.grid(*receivers, p["nz"], p["nx"], p["nts"])]
        self.wavelet = np.ascontiguousarray(wavelet, np.float32).reshape(1, -1)
        self.nExt = nExt
        self.sources = {}
        nzx = p["nz"] * p["nx"]
        traces = 4 * p["nts"] * (self.rec[0].nDev + 8)
        # engine bytes per shot by kind: wavefield pairs, the Born forward's
        # lockstep ring, the adjoints' nts source slices, offset ring and image
        self.bytes_per_shot = {"nl_forward": 4 * nzx * 2 + traces,
                               "born_forward": 4 * nzx * (2 + (5 if p["sub"] == 2 else p["nts"])) + traces,
                               "born_adjoint": 4 * nzx * (2 + p["nts"] + 2 + 1) + traces,
                               "bornext_adjoint": 4 * nzx * (2 + p["nts"] + 17 + nExt) + traces}

    def source(self, iz, ix):
        key = (iz, ix)
        if key not in self.sources:
            p = self.params
            self.sources[key] = A.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"])
        return self.sources[key]


class Server:
    def __init__(self, max_models=4, hbm_budget_gb=200.0, max_shots=32, nExt=21, device=0):
        self.max_models = max_models
        self.budget = hbm_budget_gb * 1e9
        self.max_shots = max_shots
        self.nExt = nExt
        self.device = device
        self.models = collections.OrderedDict()
        self.queue = collections.deque()
        self.calls = 0

    # -- models
    def add_model(self, model_id, params, vel, receivers, wavelet):
        """params: aiso_fd_params fields of the padded grid; vel [nx][nz];
        receivers: (nzDev, ozDev, dzDev, nxDev, oxDev, dxDev) on that grid;
        wavelet: [nts] source signature."""
        if model_id in self.models:
            self.models.move_to_end(model_id)
            return
        while len(self.models) >= self.max_models:
            _, old = self.models.popitem(last=False)
            old.ctx.close()
        self.models[model_id] = _Model(params, vel, receivers, wavelet, self.nExt, self.device)

    # -- requests
    def submit(self, req: Request):
        if req.kind not in KINDS:
            raise ValueError(f"unknown request kind {req.kind}")
        if req.model not in self.models:
            raise KeyError(f"velocity model {req.model!r} is not resident")
        self.queue.append(req)

    def pending(self):
        return len(self.queue)

    @staticmethod
    def _key(r):
        return (r.model, r.kind, r.reflectivity_id if r.kind == "born_forward" else None)

    def step(self, clock=time.perf_counter):
        """Run one batch; returns the requests it completed, their `done` set
        from `clock` once the batch has finished."""
        if not self.queue:
            return []
        key = self._key(self.queue[0])
        m = self.models[key[0]]
        self.models.move_to_end(key[0])
        # the budget is shared by the resident models; the engine itself also
        # splits a call into launches that fit the free HBM
        cap = max(1, min(self.max_shots, int(self.budget / self.max_models // m.bytes_per_shot[key[1]])))
        batch = [r for r in self.queue if self._key(r) == key][:cap]
        taken = set(map(id, batch))
        self.queue = collections.deque(r for r in self.queue if id(r) not in taken)
        self._run(m, batch)
        t = clock()
        for r in batch:
            r.done = t
        self.calls += 1
        return batch

    def _run(self, m, batch):
        p = m.params
        n, nrec, nts = len(batch), m.rec[0].nDev, p["nts"]
        srcs = [m.source(*r.source) for r in batch]
        kind = batch[0].kind
        if kind == "nl_forward":
            sig = np.repeat(m.wavelet[None], n, axis=0)                  # [nShot][1][nts]
            d = np.zeros((n, nrec, nts), np.float32)
            m.ctx.nl_forward(sig, d, srcs, m.rec)
            out = list(d)
        elif kind == "born_forward":
            d = np.zeros((n, nrec, nts), np.float32)
            m.ctx.born_forward(np.ascontiguousarray(batch[0].payload, np.float32), d, srcs, m.wavelet, m.rec)
            out = list(d)
        else:
            data = np.ascontiguousarray(np.stack([r.payload for r in batch]), np.float32)
            ext = kind == "bornext_adjoint"
            img = np.zeros((n, m.nExt if ext else 1, p["nx"], p["nz"]), np.float32)
            (m.ctx.bornext_adjoint if ext else m.ctx.born_adjoint)(img, data, srcs, m.wavelet, m.rec)
            out = list(img if ext else img[:, 0])
        for r, o in zip(batch, out):
            r.result = o


def run_stream(server: Server, requests, clock=time.perf_counter):
    """Serve `requests` (with `arrival` offsets in seconds from the start, in
    order) in real time: each is submitted once its arrival time has passed;
    batches run while requests are pending.  Returns the wall time from the
    first arrival to the last completion; each request gets `arrival` and
    `done` on the clock's scale."""
    t0 = clock()
    todo = collections.deque(sorted(requests, key=lambda r: r.arrival))
    for r in todo:
        r.arrival += t0
    while todo or server.pending():
        now = clock()
        while todo and todo[0].arrival <= now:
            server.submit(todo.popleft())
        if server.pending():
            server.step(clock)
        elif todo:
            time.sleep(max(0.0, min(todo[0].arrival - clock(), 0.01)))
    last = max(r.done for r in requests)
    return last - t0

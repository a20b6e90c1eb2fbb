"""Request server over the engine: continuous batching of one-shot modeling
and imaging requests from many callers on resident velocity models.

Request kinds, each one shot with a point source on a resident velocity
model's grid and that model's receiver line:

* ``nl_forward``: nonlinear data of the model's wavelet;
* ``born_forward``: Born data of a reflectivity (requests that name the same
  ``reflectivity_id`` share one call; without an id, requests share a call
  only when they carry the same array object);
* ``born_adjoint``: the shot's Born image of its data;
* ``bornext_adjoint``: the shot's subsurface-offset extended image
  (``nExt`` offsets, 21 by default) of its data.

Scheduling: continuous batching over slots.  Up to ``slots`` engine calls run
at once, each on its own context and HIP stream, so their kernels overlap on
the device.  Whenever a slot is free, the oldest pending request picks the
batch key (model, kind, reflectivity); every pending request with that key
joins it, up to ``max_shots`` and up to what the HBM budget admits (the
engine's bytes per shot of that kind on that model times the shots, see
below).  A request that arrives while calls run is admitted as soon as a
slot and memory free up: it does not wait for the running calls to finish.
When nothing runs, the oldest request is admitted even if it alone exceeds
the budget (the engine itself splits a call into launches that fit the free
HBM).  The adjoint kinds run with the engine's ``imagePerShot``
option, so each request gets its own image from the shared call; every
result equals what its shot gives alone.

Each resident model keeps the contexts its calls have used (velocity and
geometry tables resident in each), at most one per slot; the budget covers
what all contexts hold (each counted at the largest batch it has run), so a
batch that would grow a context's buffers beyond the budget waits until
models are evicted or, when nothing else runs, runs alone.  At most
``max_models`` models are resident; loading another evicts the least
recently used model that has no pending or running request, and frees its
contexts.

``step()`` runs one batch on the caller's thread; ``start()`` starts
``slots`` worker threads that take batches as slots free up (engine calls
release the GIL), ``drain()`` waits for the queue to empty, ``stop()`` ends
the workers.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
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
    admitted: float | None = None
    done: float | None = None
    result: np.ndarray | None = None
    error: str | None = None


class _Slot:
    """One engine context of a model, and the HBM its buffers hold (the
    largest batch it has run, by the per-shot estimate)."""

    def __init__(self):
        self.ctx = None                    # made on first use, outside the server's lock
        self.reserved = 0.0
        self.busy = False


class _Model:
    def __init__(self, params, vel, receivers, wavelet, nExt, device, options=None):
        p = dict(params, nExt=nExt, extension=A.EXT_OFFSET)
        self.params = p
        self.device = device
        self.vel = np.ascontiguousarray(vel, np.float32)
        self.rec = [A.Device.grid(*receivers, p["nz"], p["nx"], p["nts"])]
        self.wavelet = np.ascontiguousarray(wavelet, np.float32).reshape(1, -1)
        self.nExt = nExt
        self.options = dict(options or {})
        self.sources = {}
        first = _Slot()
        first.ctx = self.new_context()     # the velocity checks run when the model is added
        self.slots = [first]
        self.active = 0                    # calls running on this model
        nzx = p["nz"] * p["nx"]
        traces = 4 * p["nts"] * (self.rec[0].nDev + 8)
        # engine bytes per shot by kind: wavefield pairs, the Born forward's
        # lockstep ring, the adjoints' nts source slices, offset ring and image
        self.bytes_per_shot = {"nl_forward": 4 * nzx * 2 + traces,
                               "born_forward": 4 * nzx * (2 + (5 if p["sub"] == 2 else p["nts"])) + traces,
                               "born_adjoint": 4 * nzx * (2 + p["nts"] + 2 + 1) + traces,
                               "bornext_adjoint": 4 * nzx * (2 + p["nts"] + 17 + nExt) + traces}

    def new_context(self):
        ctx = A.Context(self.params, self.device)
        ctx.set_velocity(self.vel)
        ctx.set_option("imagePerShot", 1)
        for k, v in self.options.items():
            ctx.set_option(k, v)
        return ctx

    def source(self, iz, ix):
        key = (iz, ix)
        if key not in self.sources:
            p = self.params
            self.sources[key] = A.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"])
        return self.sources[key]

    def close(self):
        for sl in self.slots:
            if sl.ctx is not None:
                sl.ctx.close()
        self.slots = []


class Server:
    def __init__(self, max_models=4, hbm_budget_gb=200.0, max_shots=32, nExt=21, device=0, slots=4, options=None):
        self.max_models = max_models
        self.budget = hbm_budget_gb * 1e9
        self.max_shots = max_shots
        self.nExt = nExt
        self.device = device
        self.slots = slots
        self.options = dict(options or {})   # engine options (aiso_ctx_set_option) of every context
        self.models = collections.OrderedDict()
        self.queue = collections.deque()
        self.cv = threading.Condition()
        self.resident_bytes = 0.0         # HBM the models' contexts hold, by the per-shot estimates
        self.running = 0
        self.calls = 0
        self.max_running = 0               # most calls seen running at once
        self.workers = []
        self.stopping = False
        self.clock = time.perf_counter

    # -- models
    def add_model(self, model_id, params, vel, receivers, wavelet):
        """params: aiso_fd_params fields of the padded grid; vel [nx][nz];
        receivers: (nzDev, ozDev, dzDev, nxDev, oxDev, dxDev) on that grid;
        wavelet: [nts] source signature."""
        with self.cv:
            if model_id in self.models:
                self.models.move_to_end(model_id)
                return
            while len(self.models) >= self.max_models:
                busy = {r.model for r in self.queue}
                victim = next((k for k, m in self.models.items() if m.active == 0 and k not in busy), None)
                if victim is None:
                    raise RuntimeError("every resident model has pending or running requests; "
                                       "none can be evicted")
                old = self.models.pop(victim)
                self.resident_bytes -= sum(sl.reserved for sl in old.slots)
                old.close()
        m = _Model(params, vel, receivers, wavelet, self.nExt, self.device, self.options)
        with self.cv:
            self.models[model_id] = m

    # -- requests
    def submit(self, req: Request):
        if req.kind not in KINDS:
            raise ValueError(f"unknown request kind {req.kind}")
        with self.cv:
            if req.model not in self.models:
                raise KeyError(f"velocity model {req.model!r} is not resident")
            if req.kind == "born_forward" and req.payload is None:
                raise ValueError("born_forward needs a reflectivity payload")
            self.queue.append(req)
            self.cv.notify_all()

    def pending(self):
        with self.cv:
            return len(self.queue)

    @staticmethod
    def _key(r):
        if r.kind != "born_forward":
            return (r.model, r.kind, None)
        return (r.model, r.kind, ("id", r.reflectivity_id) if r.reflectivity_id is not None else ("array", id(r.payload)))

    def _take(self):
        """Under the lock: the next batch the budget admits, removed from the
        queue, with the model context (slot) it runs on, or None.  A batch
        runs on an idle context of its model (the one holding the most
        memory) or on a new one while the model has fewer than `slots`; the
        growth of that context's buffers must fit in the budget left by all
        contexts' holdings."""
        if not self.queue or self.running >= self.slots:
            return None
        key = self._key(self.queue[0])
        m = self.models[key[0]]
        idle = [sl for sl in m.slots if not sl.busy]
        if idle:
            slot = max(idle, key=lambda sl: sl.reserved)
        elif len(m.slots) < self.slots:
            slot = _Slot()
        else:
            return None
        bps = m.bytes_per_shot[key[1]]
        room = self.budget - self.resident_bytes + slot.reserved
        cap = min(self.max_shots, int(room // bps))
        if cap < 1:
            if self.running:
                return None                # wait for memory to free up
            cap = 1
        batch = [r for r in self.queue if self._key(r) == key][:cap]
        taken = set(map(id, batch))
        self.queue = collections.deque(r for r in self.queue if id(r) not in taken)
        need = bps * len(batch)
        if need > slot.reserved:
            self.resident_bytes += need - slot.reserved
            slot.reserved = need
        if slot not in m.slots:
            m.slots.append(slot)
        slot.busy = True
        self.running += 1
        self.max_running = max(self.max_running, self.running)
        m.active += 1
        self.models.move_to_end(key[0])
        t = self.clock()
        for r in batch:
            r.admitted = t
        return m, slot, batch

    def _execute(self, item):
        m, slot, batch = item
        err = None
        try:
            if slot.ctx is None:
                slot.ctx = m.new_context()
            self._run(m, slot.ctx, batch)
        except Exception as e:             # the requests carry the error; the server keeps serving
            err = f"{type(e).__name__}: {e}"
        t = self.clock()
        with self.cv:
            slot.busy = False
            self.running -= 1
            m.active -= 1
            self.calls += 1
            for r in batch:
                r.done = t
                r.error = err
            self.cv.notify_all()
        return batch

    def step(self, clock=None):
        """Run one batch on the caller's thread; returns the requests it
        completed, their `done` set from the clock once the call has ended."""
        if clock is not None:
            self.clock = clock
        with self.cv:
            item = self._take()
        return self._execute(item) if item else []

    # -- worker threads
    def start(self, clock=None):
        if clock is not None:
            self.clock = clock
        if self.workers:
            return
        self.stopping = False
        for k in range(self.slots):
            t = threading.Thread(target=self._worker, name=f"aiso-serve-{k}", daemon=True)
            t.start()
            self.workers.append(t)

    def _worker(self):
        while True:
            with self.cv:
                item = self._take()
                while item is None:
                    if self.stopping:
                        return
                    self.cv.wait()
                    item = self._take()
            self._execute(item)

    def drain(self):
        with self.cv:
            while self.queue or self.running:
                self.cv.wait()

    def stop(self):
        with self.cv:
            self.stopping = True
            self.cv.notify_all()
        for t in self.workers:
            t.join()
        self.workers = []

    def close(self):
        """Stop the workers and free every model's contexts."""
        self.stop()
        with self.cv:
            for m in self.models.values():
                m.close()
            self.models.clear()
            self.resident_bytes = 0.0

    def _run(self, m, ctx, batch):
        p = m.params
        n, nrec, nts = len(batch), m.rec[0].nDev, p["nts"]
        srcs = [m.source(*r.source) for r in batch]
        kind = batch[0].kind
        if kind == "nl_forward":
            sig = np.repeat(m.wavelet[None], n, axis=0)                  # [nShot][1][nts]
            d = np.zeros((n, nrec, nts), np.float32)
            ctx.nl_forward(sig, d, srcs, m.rec)
            out = list(d)
        elif kind == "born_forward":
            d = np.zeros((n, nrec, nts), np.float32)
            ctx.born_forward(np.ascontiguousarray(batch[0].payload, np.float32), d, srcs, m.wavelet, m.rec)
            out = list(d)
        else:
            data = np.ascontiguousarray(np.stack([r.payload for r in batch]), np.float32)
            ext = kind == "bornext_adjoint"
            img = np.zeros((n, m.nExt if ext else 1, p["nx"], p["nz"]), np.float32)
            (ctx.bornext_adjoint if ext else ctx.born_adjoint)(img, data, srcs, m.wavelet, m.rec)
            out = list(img if ext else img[:, 0])
        for r, o in zip(batch, out):
            r.result = o


def run_stream(server: Server, requests, clock=time.perf_counter):
    """Serve `requests` (with `arrival` offsets in seconds from the start, in
    order) in real time: each is submitted once its arrival time has passed.
    With worker threads started, batches run as slots free up, concurrently
    with later arrivals; otherwise batches run one at a time between
    arrivals.  Returns the wall time from the start to the last completion;
    each request gets `arrival` and `done` on the clock's scale."""
    server.clock = clock
    t0 = clock()
    todo = collections.deque(sorted(requests, key=lambda r: r.arrival))
    for r in todo:
        r.arrival += t0
    if server.workers:
        while todo:
            now = clock()
            while todo and todo[0].arrival <= now:
                server.submit(todo.popleft())
            if todo:
                time.sleep(max(0.0, min(todo[0].arrival - clock(), 0.005)))
        server.drain()
    else:
        while todo or server.pending():
            now = clock()
            while todo and todo[0].arrival <= now:
                server.submit(todo.popleft())
            if server.pending():
                server.step()
            elif todo:
                time.sleep(max(0.0, min(todo[0].arrival - clock(), 0.01)))
    last = max(r.done for r in requests)
    return last - t0

"""Differentiable nonlinear modeling for full-waveform inversion (PyTorch).

``NonlinearVelocityOperator`` maps a velocity model (padded grid [nx][nz],
float32 tensor) to the data of the nonlinear operator [nShot][nRec][nts]
through ``torch.autograd``:

* forward: ``aiso_nl_forward`` with the tensor as the context's velocity;
* backward: ``aiso_nl_velocity_gradient``, the adjoint-state gradient with
  checkpointed forward recomputation, given the incoming dL/d(data).

The reference's FWI drives ``nonlinearFwiPropShotsGpu`` with pySolver and
takes the Born adjoint as the gradient (PY/fwiMain.py:123-142,
PY/Acoustic_iso_float.py:702-806); here the gradient is the exact transpose
of the float forward recursion (sponge, source scaling and all), so any loss
written in PyTorch on the data differentiates through it.

Tensors may live on ``cuda:0`` (device pointers are handed to the engine) or
on the CPU.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _capi as A


def _device(spec, nz, nx, nts):
    if isinstance(spec, A.Device):
        return spec
    kind, args = spec
    if kind == "grid":
        return A.Device.grid(*args, nz, nx, nts)
    if kind == "ints":
        return A.Device.ints(args[0], args[1], nz, nx, nts)
    if kind == "coords":
        return A.Device.coords(*args[:2], nz, nx, *args[2:6], nts)
    raise ValueError(f"unknown device kind {kind}")


class _NlVelocity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel, op):
        mem = A.MEM_DEVICE if vel.is_cuda else A.MEM_HOST
        v = vel.detach().contiguous()
        op.context.set_velocity(v, mem=mem)
        data = torch.zeros((op.nshot, op.nrec, op.nts), dtype=torch.float32, device=vel.device)
        op.context.nl_forward(op.model_on(vel.device), data, op.sources, op.receivers, mem=mem)
        ctx.op = op
        ctx.save_for_backward(v)
        return data

    @staticmethod
    def backward(ctx, grad_data):
        (v,) = ctx.saved_tensors
        op = ctx.op
        mem = A.MEM_DEVICE if v.is_cuda else A.MEM_HOST
        op.context.set_velocity(v, mem=mem)
        g = torch.empty_like(v)
        op.context.nl_velocity_gradient(op.model_on(v.device), grad_data.detach().contiguous(), g, op.sources,
                                        op.receivers, checkpoint_every=op.checkpoint_every, mem=mem)
        return g, None


class NonlinearVelocityOperator:
    """Velocity [nx][nz] -> nonlinear data [nShot][nRec][nts], differentiable.

    params: aiso_fd_params fields (nz, nx, dz, dx, nts, dts, sub, pads, ...);
    sources / receivers: per-shot device specs (``("grid", (nzDev, ozDev,
    dzDev, nxDev, oxDev, dxDev))``, ``("ints", (z, x))``, ``("coords", (z, x,
    oz, dz, ox, dx))``) or ``_capi.Device`` objects, receivers 1 or nShot;
    wavelet: source signals [1 or nShot][nSrcDev][nts]."""

    def __init__(self, params, sources, receivers, wavelet, checkpoint_every=0):
        self.params = dict(params)
        self.context = A.Context(self.params)
        nz, nx, nts = self.params["nz"], self.params["nx"], self.params["nts"]
        self.sources = [_device(s, nz, nx, nts) for s in sources]
        self.receivers = [_device(r, nz, nx, nts) for r in receivers]
        self.model = np.ascontiguousarray(wavelet, np.float32)
        if self.model.ndim != 3:
            raise ValueError("wavelet must be [1 or nShot][nSrcDev][nts]")
        self._models = {}
        self.nshot = len(self.sources)
        self.nrec = self.receivers[0].nDev
        self.nts = nts
        self.checkpoint_every = int(checkpoint_every)

    def model_on(self, device):
        if device.type == "cpu":
            return self.model
        if device not in self._models:
            self._models[device] = torch.from_numpy(self.model).to(device)
        return self._models[device]

    def __call__(self, vel: torch.Tensor) -> torch.Tensor:
        if vel.dtype != torch.float32 or vel.shape != (self.params["nx"], self.params["nz"]):
            raise ValueError("velocity must be float32 [nx][nz] on the padded grid")
        return _NlVelocity.apply(vel, self)

    def misfit_and_gradient(self, vel: torch.Tensor, observed: torch.Tensor):
        """0.5 |F(vel) - observed|^2 and its gradient with respect to vel."""
        v = vel.detach().clone().requires_grad_(True)
        r = self(v) - observed
        loss = 0.5 * (r * r).sum()
        loss.backward()
        return loss.detach(), v.grad

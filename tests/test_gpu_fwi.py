"""Velocity gradient of the nonlinear operator on the GPU (adjoint state with
checkpointed recomputation, aiso_nl_velocity_gradient) against the oracle's,
and its torch.autograd wrapper."""
import numpy as np
import pytest

from cases import make_case, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _engine(c):
    from acoustic_isotropic_operators_amd import _capi as A
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=c.pad,
                  zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0)
    ctx = A.Context(params)
    ctx.set_velocity(c.vel)
    srcs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.src_args]
    recs = [A.Device.grid(*a, c.nz, c.nx, c.nts) for _, a in c.rec_args]
    return A, ctx, srcs, recs


@pytest.mark.parametrize("sub", [2, 3])
def test_velocity_gradient_matches_oracle(sub):
    c = make_case(seed=51, nShot=3, nts=40, sub=sub, dts=0.001 * sub)
    A, ctx, srcs, recs = _engine(c)
    m = c.wavelet[None, None, :].astype(np.float32)
    rng = np.random.default_rng(52)
    dgrad = rng.standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)
    grad = np.zeros((c.nx, c.nz), np.float32)
    ctx.nl_velocity_gradient(m, dgrad, grad, srcs, recs)
    ref = O.nl_velocity_gradient(c.grid, c.vel, c.oracle_sources(), c.oracle_receivers(), m, dgrad, "f")
    refd = O.nl_velocity_gradient(c.grid, c.vel, c.oracle_sources(), c.oracle_receivers(), m.astype(np.float64),
                                  dgrad.astype(np.float64), "d")
    print(f"gradient: relL2 vs float oracle {rel_l2(grad, ref):.3e} (bitwise equal: {np.array_equal(grad, ref)}), "
          f"vs FP64 oracle {rel_l2(grad, refd):.3e}")
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(grad, ref)


def test_checkpoint_interval_does_not_change_the_gradient():
    c = make_case(seed=53, nShot=2, nts=30, sub=2, dts=0.002)
    A, ctx, srcs, recs = _engine(c)
    m = c.wavelet[None, None, :].astype(np.float32)
    dgrad = np.random.default_rng(54).standard_normal((c.nShot, recs[0].nDev, c.nts)).astype(np.float32)
    outs = []
    T = (c.nts - 1) * c.sub
    for every in (0, 1, 5, T):
        g = np.zeros((c.nx, c.nz), np.float32)
        ctx.nl_velocity_gradient(m, dgrad, g, srcs, recs, checkpoint_every=every)
        outs.append(g)
    for g in outs[1:]:
        np.testing.assert_array_equal(g, outs[0])


def test_gradient_against_fp64_finite_differences():
    """The engine's float gradient against central differences of the misfit
    evaluated in FP64 (the oracle's forward), along smooth directions."""
    c = make_case(seed=55, nShot=2, nts=40, sub=2, dts=0.002)
    A, ctx, srcs, recs = _engine(c)
    g, os_, or_ = c.grid, c.oracle_sources(), c.oracle_receivers()
    m = c.wavelet[None, None, :].astype(np.float32)
    zz, xx = np.meshgrid(np.arange(c.nz), np.arange(c.nx))
    vtrue = c.vel + 60.0 * np.exp(-((zz - 30) ** 2 + (xx - 40) ** 2) / 50.0)
    dobs = O.nl_forward(g, vtrue, os_, or_, m.astype(np.float64), "d")

    def misfit(v):
        r = O.nl_forward(g, v, os_, or_, m.astype(np.float64), "d") - dobs
        return 0.5 * float(np.sum(r * r))

    v0 = c.vel.astype(np.float64)
    # The residual is ~1e-4 of the data here, so it is formed in FP64 and
    # rounded once: the test then measures the gradient operator, not the
    # cancellation of two float forward runs.
    r = (O.nl_forward(g, v0, os_, or_, m.astype(np.float64), "d") - dobs).astype(np.float32)
    grad = np.zeros((c.nx, c.nz), np.float32)
    ctx.nl_velocity_gradient(m, r, grad, srcs, recs)
    rng = np.random.default_rng(56)
    for k in range(3):
        dv = rng.standard_normal((c.nx, c.nz))
        kern = np.ones(9) / 9.0
        for _ in range(2):
            dv = np.apply_along_axis(lambda a: np.convolve(a, kern, mode="same"), 0, dv)
            dv = np.apply_along_axis(lambda a: np.convolve(a, kern, mode="same"), 1, dv)
        dv /= np.abs(dv).max()
        eps = 1e-2
        fd = (misfit(v0 + eps * dv) - misfit(v0 - eps * dv)) / (2 * eps)
        an = float(np.sum(grad.astype(np.float64) * dv))
        print(f"direction {k}: FP64 finite difference {fd:.8e}, engine gradient {an:.8e}, rel {abs(fd - an) / abs(fd):.2e}")
        assert abs(fd - an) <= 1e-4 * abs(fd)


@pytest.mark.parametrize("where", ["cuda:0", "cpu"])
def test_torch_autograd_function(where):
    """loss.backward() through the operator gives the engine's gradient for
    dL/d(data) = residual; tensors on the GPU or on the host."""
    import torch

    from acoustic_isotropic_operators_amd.fwi import NonlinearVelocityOperator
    c = make_case(seed=57, nShot=2, nts=40, sub=2, dts=0.002)
    params = dict(nz=c.nz, nx=c.nx, dz=c.dz, dx=c.dx, nts=c.nts, dts=c.dts, sub=c.sub, zPadMinus=c.pad,
                  zPadPlus=c.pad, xPadMinus=c.pad, xPadPlus=c.pad, alphaCos=c.alpha, fMax=20.0)
    op = NonlinearVelocityOperator(params, [("grid", a) for _, a in c.src_args], [("grid", a) for _, a in c.rec_args],
                                   c.wavelet[None, None, :])
    dev = torch.device(where)
    v = torch.tensor(c.vel, device=dev, requires_grad=True)
    dobs = torch.zeros((c.nShot, op.nrec, c.nts), device=dev)
    d = op(v)
    loss = 0.5 * ((d - dobs) ** 2).sum()
    loss.backward()
    r = (d - dobs).detach().cpu().numpy()
    ref = O.nl_velocity_gradient(c.grid, c.vel, c.oracle_sources(), c.oracle_receivers(),
                                 c.wavelet[None, None, :].astype(np.float32), r, "f")
    np.testing.assert_array_equal(v.grad.cpu().numpy(), ref)
    of = O.nl_forward(c.grid, c.vel, c.oracle_sources(), c.oracle_receivers(), c.wavelet[None, None, :], "f")
    np.testing.assert_array_equal(d.detach().cpu().numpy(), of)


def test_engine_gradient_close_to_engine_born_adjoint():
    """Same comparison on the GPU: the engine's exact gradient against the
    engine's Born adjoint of the same residual (cosine >= 0.99)."""
    import pyAcoustic_iso_float_born as born
    import pyAcoustic_iso_float_nl as nl
    c = make_case(seed=58, nShot=3, nts=80, sub=2, dts=0.002)
    A, ctx, srcs, recs = _engine(c)
    m = c.wavelet[None, None, :].astype(np.float32)
    zz, xx = np.meshgrid(np.arange(c.nz), np.arange(c.nx))
    ctx.set_velocity((c.vel + 80.0 * np.exp(-((zz - 30) ** 2 + (xx - 40) ** 2) / 72.0)).astype(np.float32))
    dobs = np.zeros((c.nShot, recs[0].nDev, c.nts), np.float32)
    ctx.nl_forward(m, dobs, srcs, recs)
    ctx.set_velocity(c.vel)
    d = np.zeros_like(dobs)
    ctx.nl_forward(m, d, srcs, recs)
    r = d - dobs
    gn = np.zeros((c.nx, c.nz), np.float32)
    ctx.nl_velocity_gradient(m, r, gn, srcs, recs)
    op = born.BornShotsGpu(c.vel_vector(), c.par, c.product_sources(nl), [c.wavelet[None, :].copy()],
                           c.product_receivers(nl))
    gb = np.zeros((c.nx, c.nz), np.float32)
    op.adjoint(False, gb, r)
    a, b = gn.ravel().astype(np.float64), gb.ravel().astype(np.float64)
    cos = float(a @ b / np.linalg.norm(a) / np.linalg.norm(b))
    print(f"cosine {cos:.4f}")
    assert cos >= 0.99

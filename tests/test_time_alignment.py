"""Absolute time alignment of injection and recording, by closed form.

With sub = 1 (fine step = coarse step, interpolation filter f = [1, 0 | 0, 1]),
a delta wavelet w = delta(t) and a receiver on the source point, the
reference's nonlinear forward (F/src/nonlinearShotsGpuFunctions.cu:278-287:
step, injectSource of S[itw - 1], damp, recordInterpData into its and its+1)
gives, away from the sponge:

    d[0] = 0,
    d[1] = S[0] = (dtw * dtw * v * v) * w[0]                    (scaleSeismicSource)
    d[2] = (v2dt2 * ((zc0 * S0) + (xc0 * S0)) + S0) + S0         (one stencil step)

and its adjoint (:620-630: step, interpInjectData of d[its]*f[it2+1] +
d[its+1]*f[h+it2+1], damp, recordSource at itw) maps d = delta at coarse
sample 1 to the model sample 0 only, with the same scale.  A one-step shift
of the injection or the record, in the forward or the adjoint, moves the
nonzero sample and fails these equalities; the relations are the
mathematics of the scheme, not a restatement of the oracle's code.
"""
import numpy as np
import pytest

from cases import make_case
from oracle import oracle as O

IZ, IX = 29, 37


def _case():
    return make_case(seed=91, nz=58, nx=74, nts=20, dts=0.001, sub=1, pad=10, nShot=1)


def _expected(c):
    v = np.float32(c.vel[IX, IZ])
    dtw = np.float32(c.dts) / np.float32(c.sub)
    sc = np.float32(np.float32(np.float32(dtw * dtw) * v) * v)
    v2 = np.float32(np.float32(np.float32(v * v) * dtw) * dtw)
    C = np.array([-2.927222222, 1.666666667, -0.238095238, 0.039682539, -0.004960317, 0.000317460])
    zc0 = np.float32(C[0] / (c.dz * c.dz))
    xc0 = np.float32(C[0] / (c.dx * c.dx))
    lap = np.float32(np.float32(zc0 * sc) + np.float32(xc0 * sc))
    d2 = np.float32(np.float32(np.float32(v2 * lap) + sc) + sc)
    return sc, d2


def _delta(c):
    w = np.zeros((1, 1, c.nts), np.float32)
    w[0, 0, 0] = 1.0
    return w


def _check_forward(c, d):
    sc, d2 = _expected(c)
    tr = d[0, 0]
    print(f"forward trace at the source: {tr[:4]}, expected [0, {sc}, {d2}, ...]")
    assert tr[0] == 0.0
    assert tr[1] == sc
    assert abs(float(tr[2]) - float(d2)) <= 2e-6 * abs(float(d2))


def _check_adjoint(c, m):
    sc, _ = _expected(c)
    print(f"adjoint model samples: {m[0, 0, :4]}, expected [{sc}, 0, 0, ...]")
    assert m[0, 0, 0] == sc
    assert not np.any(m[0, 0, 1:])


def _data_delta(c):
    d = np.zeros((1, 1, c.nts), np.float32)
    d[0, 0, 1] = 1.0
    return d


def test_oracle_time_alignment():
    c = _case()
    g = c.grid
    src = [O.Device.ints([IZ], [IX], c.nz, c.nx, c.nts)]
    rec = [O.Device.ints([IZ], [IX], c.nz, c.nx, c.nts)]
    d = O.nl_forward(g, c.vel, src, rec, _delta(c), "f")
    _check_forward(c, d)
    m = O.nl_adjoint(g, c.vel, src, rec, _data_delta(c), 1, "f")
    _check_adjoint(c, m)


@pytest.mark.gpu
def test_engine_time_alignment():
    import pyAcoustic_iso_float_nl as nl
    c = _case()
    v = c.vel_vector()
    src = [nl.deviceGpu([IZ], [IX], v, c.nts, 0, 0, 0)]
    rec = [nl.deviceGpu([IZ], [IX], v, c.nts, 0, 0, 0)]
    op = nl.nonlinearPropShotsGpu(v, c.par, src, rec)
    d = np.zeros((1, 1, c.nts), np.float32)
    op.forward(False, _delta(c), d)
    _check_forward(c, d)
    m = np.zeros((1, 1, c.nts), np.float32)
    op.adjoint(False, m, _data_delta(c))
    _check_adjoint(c, m)

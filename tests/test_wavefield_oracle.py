"""The oracle's saved wavefields (forwardWavefield / adjointWavefield of the
reference), pinned by what the scheme fixes rather than by its code:

* recordInterpData and interpWavefield apply the same weights to the same
  fine slices (F/src/kernelsGpu.cu:119-141), so a forward wavefield sampled
  at the receiver points equals the regular-grid data traces, bit for bit:
  the nonlinear wavefield against the nonlinear data, the Born source
  wavefield against the nonlinear data of the Born source signal, the Born
  scattered wavefield against the Born data;
* the adjoint wavefield at a source point is the time-interpolation adjoint
  of the source trace the adjoint records there (recordSource, :113-116);
* the Born adjoint's scaled receiver wavefield imaged against the source
  wavefield, sum_k W[k] * sec[k], is the Born image (scaling applied per
  sample instead of once: equal to float rounding).
"""
import numpy as np

from cases import make_case, rel_l2
from oracle import oracle as O


def _case(sub):
    return make_case(seed=61, nShot=2, sub=sub, dts=0.001 * sub, nts=40)


def _at(wf, pos, nz):
    pos = np.asarray(pos)
    return wf[:, pos // nz, pos % nz].T          # [point][nts]


def test_nl_forward_wavefield_at_receivers_is_the_data():
    for sub in (2, 3):
        c = _case(sub)
        g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
        m = c.wavelet[None, None, :].astype(np.float32)
        wf = O.nl_forward_wavefield(g, c.vel, src, rec, m, 1)
        v2, _ = O.vel_scales(g, c.vel)
        reg = src[1].adjoint(m[0])
        O.scale_source(g, c.vel, src[1].regPos, reg)
        fine = O.interp_fwd(g, reg)
        dreg = O.nl_fwd_reg(g, v2, fine, src[1].regPos, rec[0].regPos)
        assert np.abs(dreg).max() > 0
        np.testing.assert_array_equal(_at(wf, rec[0].regPos, c.nz), dreg)


def test_born_wavefields_at_receivers_are_the_data():
    c = _case(2)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    wav = c.wavelet[None, :].copy()
    m = np.random.default_rng(62).standard_normal((c.nx, c.nz)).astype(np.float32)
    W, Wsec = O.born_wavefields(g, c.vel, src, [wav], rec, m, 0, adjoint=False)
    v2, rf = O.vel_scales(g, c.vel)
    fine = O.born_source_fine(g, c.vel, src[0], wav)
    np.testing.assert_array_equal(_at(W, rec[0].regPos, c.nz), O.nl_fwd_reg(g, v2, fine, src[0].regPos, rec[0].regPos))
    dreg = O.born_fwd_reg(g, v2, rf, fine, src[0].regPos, rec[0].regPos, m)
    assert np.abs(dreg).max() > 0
    np.testing.assert_array_equal(_at(Wsec, rec[0].regPos, c.nz), dreg)
    # the saved wavefields are zero on the FAT ring, like the reference's zeroed buffers
    assert not W[:, :5].any() and not W[:, :, -5:].any() and not Wsec[:, :, :5].any()


def test_nl_adjoint_wavefield_at_the_source_is_the_recorded_trace():
    c = _case(2)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    d = np.random.default_rng(63).standard_normal((c.nShot, rec[0].nDev, c.nts)).astype(np.float32)
    wf = O.nl_adjoint_wavefield(g, c.vel, src, rec, d, 0)
    v2, _ = O.vel_scales(g, c.vel)
    sig = O.nl_adj_reg(g, v2, rec[0].adjoint(d[0]), src[0].regPos, rec[0].regPos)
    # wf[k] = sum over fine steps itw of p_itw * (weight of sample k); the
    # adjoint records p_itw at the source: the same sums over sig
    f = O.tables(g)[2]
    h = g.sub + 1
    ref = np.zeros((len(src[0].regPos), g.nts), np.float64)
    for its in range(g.nts - 1):
        for it2 in range(g.sub):
            itw = its * g.sub + it2
            ref[:, its] += sig[:, itw] * f[it2]
            ref[:, its + 1] += sig[:, itw] * f[h + it2]
    got = _at(wf, src[0].regPos, c.nz)
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) < 1e-6


def test_born_adjoint_receiver_wavefield_images_to_the_born_image():
    c = _case(2)
    g, src, rec = c.grid, c.oracle_sources(), c.oracle_receivers()
    wav = c.wavelet[None, :].copy()
    d = np.random.default_rng(64).standard_normal((c.nShot, rec[0].nDev, c.nts)).astype(np.float32)
    W, Wsec = O.born_wavefields(g, c.vel, src, [wav], rec, d, 1, adjoint=True)
    img = O.born_adjoint(g, c.vel, src[1:2], [wav], rec, d[1:2])
    got = np.einsum("kxz,kxz->xz", W.astype(np.float64), Wsec.astype(np.float64))
    assert np.abs(img).max() > 0
    print(f"sum_k W[k] sec[k] vs image: relL2 {rel_l2(got, img):.2e}")
    assert rel_l2(got, img) < 1e-5

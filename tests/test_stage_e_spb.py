"""Stage (e): sPb, gPb and the computeGPb output transforms; whole path.

Oracle pins (CPU): the 8 derivative filters are zero-mean, unit-L1 and
antisymmetric; a step edge in an eigenvector gives the strongest sPb at the
orientation whose half-disc split runs along the edge, the same orientation
at which the chi-squared cue of that edge peaks (the sPb and mPb orientation
conventions agree); the calibration sigmoid at known points.  GPU: sPb from
the same eigenpairs within 1e-5 relative; the whole path against the oracle.
"""
import numpy as np
import pytest

from conftest import synth_packed


def test_spb_filters_shape(orc):
    F = orc.spb_filters()
    for o in range(8):
        assert abs(F[o].sum()) < 1e-14
        assert abs(np.abs(F[o]).sum() - 1) < 1e-14
        np.testing.assert_allclose(F[o], -F[o][::-1, ::-1], atol=1e-12)  # odd filter


def test_spb_orientation_matches_cues(orc):
    H, W = 24, 24
    vert = np.zeros((H, W))
    vert[:, 12:] = 1.0  # edge between columns 11 and 12 (vertical line)
    ev = np.array([0.0, 0.01])
    s = orc.spb(np.stack([np.zeros((H, W)), vert]), ev)
    centre = s[:, 10, 11]
    assert centre.argmax() == 0 and centre[4] < 1e-12
    s2 = orc.spb(np.stack([np.zeros((H, W)), vert.T.copy()]), ev)
    assert s2[:, 11, 10].argmax() == 4
    lab = (vert > 0).astype(np.uint8) * 3 + 1
    cue = orc.hist_gradient(orc.mirror_u8(lab, 6), 6, 5, None)
    assert cue[:, 10, 12].argmax() == 0


def test_sigmoid_known_points(orc):
    g = np.zeros((8, 1, 3))
    g[:, 0, 0], g[:, 0, 1], g[:, 0, 2] = 0.0, 0.5, 2.0
    nm = np.array([[0.01, 0.06, 0.5]])
    gmax, gall = orc.gpb_outputs(g, nm)
    f = lambda v: (1 / (1 + np.exp(2.6433 - 10.7998 * min(1.0, 1.2 * v))) - 0.0667) / 0.9333
    np.testing.assert_allclose(gall[0, :, 0], [f(0.0), f(0.5), f(2.0)], rtol=1e-6, atol=1e-7)  # FP32 map
    np.testing.assert_allclose(gmax[0], [0.0, 0.5, 2.0], rtol=1e-7)


@pytest.mark.gpu
def test_spb_gpu_from_same_evecs(gpb, orc, ctx):
    print("spb parity 64x48", flush=True)
    rng = np.random.default_rng(1)
    K, H, W = 9, 48, 64
    vec = rng.normal(size=(K, H, W)).cumsum(1).cumsum(2)
    ev = np.sort(rng.random(K) * 0.05)
    ev[0] = 0.0
    ref = orc.spb(vec, ev)
    got = ctx.spb_from_evecs(vec.astype(np.float32), ev)
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5 * ref.max())


@pytest.mark.gpu
@pytest.mark.parametrize("W,H,seed", [(64, 48, 1), (96, 80, 2)])
def test_whole_path_gpu_vs_oracle(gpb, orc, ctx, W, H, seed):
    print(f"whole path {W}x{H} seed {seed}", flush=True)
    rgb, packed = synth_packed(gpb, W, H, regions=8, seed=seed)
    cb = gpb.make_codebook(packed, seed=seed)
    ref = orc.pipeline(packed, cb, K=9, tol=1e-10)
    ctx.set_eig_options(nvec=9, basis=40, keep=20, tol=1e-6)
    out = ctx.compute(packed, cb)
    flips = np.count_nonzero((out["mpb_nmax"] > 0) != (ref["mpb_nmax"] > 0))
    err_nm = np.abs(out["mpb_nmax"] - ref["mpb_nmax"]).max()
    err_or = np.abs(out["gpb_orient"] - ref["gpb"]).max()
    err_all = np.abs(out["gpb_all"] - ref["gpb_all"]).max()
    err_max = np.abs(out["gpb_max"] - ref["gpb_max"]).max()
    st = ctx.stats()
    print(f"  nmax flips {flips} |dnmax| {err_nm:.2e} |dgPb_o| {err_or:.2e} |dgPb_all| {err_all:.2e} "
          f"|dgPb_max| {err_max:.2e} matvecs {st['matvecs']} ms {st['stage_ms']}", flush=True)
    if flips == 0:
        assert err_or <= 1e-3 and err_all <= 1e-3 and err_max <= 1e-3

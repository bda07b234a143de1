"""Stage (e): sPb, gPb and the computeGPb output transforms; whole path.

Oracle pins (CPU): the 8 derivative filters are zero-mean, unit-L1 and
antisymmetric; a step edge in an eigenvector gives the strongest sPb at the
orientation whose half-disc split runs along the edge, the same orientation
at which the chi-squared cue of that edge peaks (the sPb and mPb orientation
conventions agree); a linear ramp's closed-form response in the interior and
at the REFLECT_101 border; the vector range 1..K-1, the 1/sqrt(lambda) scale,
the lambda <= DBL_EPSILON skip and the min-max normalisation; properties of
the calibration sigmoid.  GPU: sPb from the same eigenpairs within 1e-5
relative; the whole path against the oracle.
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


def _ramp_response(F, offsets):
    """sum_{a,b} F[a, b] * offsets[b]: a filter's correlation with an image
    that depends on the column only, through the 7 column offsets given."""
    return float((F * np.asarray(offsets, np.float64)[None, :]).sum())


def test_spb_linear_ramp_closed_form(orc):
    """One eigenvector, a column ramp stored as 5 * ramp - 2 (so a skipped
    min-max normalisation scales the answer by 5), lambda = 0.04 (1/sqrt = 5,
    1/lambda = 25).  In the interior a zero-sum filter's response to
    x / (W - 1) is sum F[a, b] (b - 3) / (W - 1); at column 0 the REFLECT_101
    border (src/spectralPb.cpp:316) makes the ramp |x| there, which differs
    from REFLECT's (the -1 -> 0 mirror) on every odd filter column."""
    H, W = 16, 21
    ramp = np.tile(np.arange(W, dtype=np.float64) / (W - 1), (H, 1))
    vec = np.stack([np.ones((H, W)), 5 * ramp - 2])
    lam = 0.04
    s = orc.spb(vec, np.array([0.0, lam]))
    F = orc.spb_filters().reshape(8, 7, 7)
    for o in range(8):
        interior = abs(_ramp_response(F[o], np.arange(-3, 4))) / (W - 1) / np.sqrt(lam)
        np.testing.assert_allclose(s[o, 3:-3, 3:-3], interior, rtol=1e-12, atol=1e-15)
        r101 = abs(_ramp_response(F[o], np.abs(np.arange(-3, 4)))) / (W - 1) / np.sqrt(lam)
        np.testing.assert_allclose(s[o, 3:-3, 0], r101, rtol=1e-12, atol=1e-15)
        r_reflect = abs(_ramp_response(F[o], [2, 1, 0, 0, 1, 2, 3])) / (W - 1) / np.sqrt(lam)
        if abs(r_reflect - r101) > 1e-6:
            assert abs(s[o, 5, 0] - r_reflect) > 1e-6
    # the ramp's direction: orientation 0 (split along columns) responds most,
    # orientation 4 not at all
    assert s[0, 8, 10] == s[:, 8, 10].max() and s[4, 8, 10] < 1e-14


def test_spb_vector_range_and_eigenvalue_skip(orc):
    """Only vectors 1 .. K-1 contribute (index 0 is the trivial pair), each
    once, scaled by 1/sqrt(lambda_i); a vector whose lambda is at most
    DBL_EPSILON contributes nothing (src/spectralPb.cpp:278-320); responses add
    in absolute value."""
    rng = np.random.default_rng(5)
    H, W = 14, 17
    a, b, c = (rng.normal(size=(H, W)).cumsum(0).cumsum(1) for _ in range(3))
    one = orc.spb(np.stack([np.zeros((H, W)), a]), np.array([0.0, 0.09]))
    two = orc.spb(np.stack([np.zeros((H, W)), b]), np.array([0.0, 0.25]))
    # index 0 ignored whatever it holds, and the last index included
    got = orc.spb(np.stack([c, a, b]), np.array([0.01, 0.09, 0.25]))
    np.testing.assert_allclose(got, one + two, rtol=1e-12, atol=1e-12 * got.max())
    # a tiny eigenvalue drops its vector; one just above DBL_EPSILON does not
    eps = np.finfo(np.float64).eps
    np.testing.assert_allclose(orc.spb(np.stack([c, a, b]), np.array([0.0, 0.09, eps])), one, rtol=1e-12,
                               atol=1e-12 * one.max())
    big = orc.spb(np.stack([c, a, b]), np.array([0.0, 0.09, 4 * eps]))
    assert big.max() > 1e6 * one.max()
    # scaling: the same vector at lambda / 4 gives twice the response
    np.testing.assert_allclose(orc.spb(np.stack([c, a]), np.array([0.0, 0.09 / 4])), 2 * one, rtol=1e-12,
                               atol=1e-12 * one.max())
    # min-max normalisation makes the response invariant to an affine map of
    # the vector with a positive scale, and a flat vector contributes nothing
    np.testing.assert_allclose(orc.spb(np.stack([c, 3 * a + 7]), np.array([0.0, 0.09])), one, rtol=1e-9,
                               atol=1e-12 * one.max())
    flat = orc.spb(np.stack([c, np.full((H, W), 2.5)]), np.array([0.0, 0.09]))
    assert np.abs(flat).max() < 1e-12


def test_sigmoid_calibration_properties(orc):
    """The computeGPb calibration (damascene/src/postprocess/skeleton.cu:229-249):
    non-decreasing; flat at and below 0 and at and above 1/1.2 (the clip of
    1.2 g to [0, 1]); 0.5 through the logistic exactly where
    10.7998 * 1.2 g = 2.6433; and the masked maximum is the largest
    orientation's value where mPb_nmax > 0.05, else 0."""
    gs = np.concatenate([[-1.0, -0.1, 0.0], np.linspace(0.0, 1.0, 41), [2.0, 5.0]])
    g = np.zeros((8, 1, gs.size))
    for o in range(8):
        g[o, 0] = gs - 0.01 * o
    g[3, 0] = gs  # orientation 3 is the largest everywhere
    nm = np.where(np.arange(gs.size) % 2 == 0, 0.06, 0.05)[None, :]
    gmax, gall = orc.gpb_outputs(g, nm)
    block3 = gall[7 - 3, :, 0]
    assert np.all(np.diff(block3) >= 0)
    assert block3[0] == block3[1] == block3[2]
    assert block3[-1] == block3[-2] == block3[np.searchsorted(gs[3:-2], 1 / 1.2) + 3]
    mid = (0.5 - 0.0667) / 0.9333
    g_mid = 2.6433 / 10.7998 / 1.2
    one = orc.gpb_outputs(np.full((8, 1, 1), g_mid), np.ones((1, 1)))[1]
    assert abs(float(one[0, 0, 0]) - mid) < 1e-6
    np.testing.assert_array_equal(gmax[0], np.where(nm[0] > 0.05, gs.astype(np.float32), 0.0))


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
    # FP32 cues may flip nonmax survival at a pixel whose oriented values tie
    # to FP32 rounding; a handful at most, and mPb_nmax within 1e-4 elsewhere
    assert flips <= 2
    keep = (out["mpb_nmax"] > 0) == (ref["mpb_nmax"] > 0)
    assert np.abs(out["mpb_nmax"] - ref["mpb_nmax"])[keep].max() <= 1e-4
    if flips == 0:
        same = (out["mpb_nmax"] > 0.05) == (ref["mpb_nmax"] > 0.05)
        assert err_or <= 1e-3 and err_all <= 1e-3
        assert np.abs(out["gpb_max"] - ref["gpb_max"])[same].max() <= 1e-3
    # stages (c)-(e) always: the oracle continued from the GPU's own mPb_nmax
    Wd, D = orc.build_w(out["mpb_nmax"])
    ev, vec, _ = orc.eigs(Wd, D, K=9, m=40, tol=1e-10)
    g = orc.gpb(ref["cues"], orc.spb(vec, ev))
    gmax, gall = orc.gpb_outputs(g, out["mpb_nmax"])
    e2 = max(np.abs(out["gpb_orient"] - g).max(), np.abs(out["gpb_all"] - gall).max(),
             np.abs(out["gpb_max"] - gmax).max())
    print(f"  against the oracle from the GPU's mPb_nmax: gPb {e2:.2e}", flush=True)
    assert e2 <= 1e-3

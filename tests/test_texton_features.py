"""Texton histogram features of the superpixel classifier
(SPFeatures::genTextonFeatures, src/SPFeatures.cpp:259-272, and the texton
part of genRecPatchFeatures, :478-545).

Oracle pins (CPU): a superpixel whose pixels all carry texton t gets 10 at t
and 0 elsewhere; one split k : n - k between two textons gets 10 k / n and
10 (n - k) / n; every row sums to 10; a window centred on a corner covers
33 x 33 pixels, one on an edge 33 x 65, an interior one 65 x 65; the features
do not depend on where a superpixel's pixels lie.  GPU: bit-exact against the
oracle on the stage (a) texton map of the same image (the Voronoi regions as
superpixels, windows at corners, edges and inside)."""
import numpy as np
import pytest

from conftest import synth_packed


def test_texton_features_closed_forms(orc):
    H, W = 12, 10
    t = np.zeros((H, W), np.uint8)
    sp = np.zeros((H, W), np.int32)
    t[:, :] = 7
    sp[:, 5:] = 1
    t[:3, 5:] = 30  # superpixel 1 (12 x 5 = 60 pixels): 15 with texton 30, 45 with texton 7
    d = orc.texton_features(t, sp, 3)
    assert d[0, 7] == 10.0 and np.count_nonzero(d[0]) == 1
    assert d[1, 30] == 15 / 60 * 10 and d[1, 7] == 45 / 60 * 10
    assert np.all(d[2] == 0)  # no pixels
    np.testing.assert_allclose(d[:2].sum(1), 10.0, rtol=1e-15)
    # pixel positions do not matter: permute pixels jointly
    rng = np.random.default_rng(1)
    perm = rng.permutation(H * W)
    d2 = orc.texton_features(t.reshape(-1)[perm].reshape(H, W), sp.reshape(-1)[perm].reshape(H, W), 3)
    np.testing.assert_array_equal(d, d2)
    with pytest.raises(ValueError):
        orc.texton_features(t, sp + 2, 3)


def test_texton_patch_features_windows(orc):
    H, W = 100, 90
    rng = np.random.default_rng(2)
    t = rng.integers(0, 32, (H, W)).astype(np.uint8)
    ctr = np.array([[0, 0], [0, 45], [50, 45], [99, 89], [40, 2]])
    d = orc.texton_patch_features(t, ctr)
    areas = [33 * 33, 33 * 65, 65 * 65, 33 * 33, 65 * 35]
    wins = [t[:33, :33], t[:33, 13:78], t[18:83, 13:78], t[67:, 57:], t[8:73, :35]]
    for k in range(5):
        np.testing.assert_allclose(d[k], np.bincount(wins[k].reshape(-1), minlength=32) / areas[k] * 10, rtol=1e-15)
        assert abs(d[k].sum() - 10) < 1e-12


@pytest.mark.gpu
def test_texton_features_gpu(gpb, orc, ctx):
    W, H = 320, 240
    print(f"texton features {W}x{H}", flush=True)
    rgb, packed = synth_packed(gpb, W, H, regions=40, seed=12)
    _, sp = gpb.synth_voronoi(W, H, 40, 6.0, 12, with_regions=True)
    cb = gpb.make_codebook(packed, seed=12)
    t = ctx.texton(packed, cb)
    got = ctx.texton_features(packed, cb, sp, 41)  # label 40 has no pixels
    np.testing.assert_array_equal(got, orc.texton_features(t, sp, 41))
    ctr = np.array([[0, 0], [0, W - 1], [H - 1, 0], [H - 1, W - 1], [120, 3], [5, 160], [120, 160], [200, 300]])
    got = ctx.texton_patch_features(packed, cb, ctr)
    np.testing.assert_array_equal(got, orc.texton_patch_features(t, ctr))
    with pytest.raises(gpb.GpbError):
        ctx.texton_patch_features(packed, cb, np.array([[H, 0]]))
    with pytest.raises(gpb.GpbError):
        ctx.texton_features(packed, cb, sp, 10)

"""The oracle's stage (b) functions against the reference's own lib_image
(globalPb/src/math/libraries/lib_image.cc and its support library, compiled
from the reference tree into oracle/_ref by `make -C oracle ref`).

Each check is bit-exact: the histogram-smoothing kernels
(lib_image::gaussian), gamma/Lab/normalise/quantise (lib_image.cc:123-245,
2315-2344), the half-sample border mirror (:308-370), the half-disc
chi-squared gradient (hist_gradient_2D, :3575-3686 with weight_matrix_disc,
orientation_slice_map, conv_in_place_1D and the X2 functor) for every radius
and kernel the cue path uses, and the oriented nonmax (nonmax_channels.cpp +
nonmax_oriented_2D + multiscalePb's clip).  The GPU kernels are in turn
bit-exact against the oracle (test_stage_b_cues.py).  Skipped where the
reference tree is absent (no _ref).
"""
import numpy as np
import pytest

from conftest import synth_packed


@pytest.fixture(scope="module")
def refl(orc):
    if not orc.REF_LIBIMAGE.exists():
        pytest.skip("reference lib_image not built (oracle/_ref needs the reference tree)")
    return orc


def test_gaussian_kernels(refl):
    for sigma in (0.1 * 25, 0.05 * 25, 0.7, 3.3):
        np.testing.assert_array_equal(refl.gaussian(sigma), refl.ref_gaussian(sigma))


def test_lab_quantize(refl, gpb):
    rng = np.random.default_rng(1)
    rgb = rng.integers(0, 256, (1 << 18, 3), dtype=np.uint8)
    corners = np.array([[a, b, c] for a in (0, 1, 254, 255) for b in (0, 128, 255) for c in (0, 2, 255)], np.uint8)
    rgb = np.concatenate([rgb, corners])
    n = rgb.shape[0]
    packed = gpb.pack_rgb(rgb.reshape(1, n, 3))
    np.testing.assert_array_equal(refl.lab_quantize(packed).reshape(3, -1), refl.ref_lab_quantize(rgb))


def test_border_mirror(refl):
    rng = np.random.default_rng(2)
    for (h, w, b) in [(9, 13, 4), (31, 37, 30), (5, 6, 3)]:
        img = rng.integers(0, 256, (h, w)).astype(np.uint8)
        np.testing.assert_array_equal(refl.mirror_u8(img, b).astype(np.float64), refl.ref_mirror(img.astype(float), b))


@pytest.mark.parametrize("r,nlab,ksigma", [(3, 25, 2.5), (5, 25, 2.5), (10, 25, 2.5), (5, 25, 1.25), (10, 25, 1.25),
                                           (20, 25, 1.25), (5, 32, 0.0), (10, 32, 0.0), (20, 32, 0.0)])
def test_hist_gradient(refl, r, nlab, ksigma):
    rng = np.random.default_rng(r * 100 + nlab)
    B = r + 2
    h, w = 2 * B + 11, 2 * B + 14
    # blocky labels (regions), with some salt so that every bin occurs
    lab = np.kron(rng.integers(0, nlab, (h // 4 + 1, w // 4 + 1)), np.ones((4, 4), np.int64))[:h, :w]
    salt = rng.random((h, w)) < 0.1
    lab[salt] = rng.integers(0, nlab, salt.sum())
    lab = lab.astype(np.uint8)
    k = refl.gaussian(ksigma) if ksigma > 0 else None
    mine = refl.hist_gradient(lab, B, r, k)
    ref = refl.ref_hist_gradient(lab, r, ksigma)[:, B:-B, B:-B]
    np.testing.assert_array_equal(mine, ref)


def test_nonmax(refl, gpb):
    rng = np.random.default_rng(3)
    mp = rng.random((8, 20, 24)) * (rng.random((8, 20, 24)) < 0.7)
    np.testing.assert_array_equal(refl.nonmax(mp), refl.ref_nonmax(mp))
    _, packed = synth_packed(gpb, 48, 40, regions=7, seed=4)
    cb = gpb.make_codebook(packed, seed=4)
    _, mpb_all = refl.local_cues(packed, refl.texton(packed, cb))
    np.testing.assert_array_equal(refl.nonmax(mpb_all), refl.ref_nonmax(mpb_all))


def damascene_block_orientations(orc):
    """Which of our orientations lands in each block of computeGPb's
    hostGPbAllConcat, derived step by step from the reference's own code:

    1. slot rule: the loop at damascene/src/spectralPb/spectralPb.cu:706-712
       writes gaussian_2D(angle pi/2 + k pi/8) into slot 7 - k;
    2. kernel indexing: spectralPb_kernel (spectralPb.cu:83-662) multiplies the
       texel at (ix + u, iy + v) by slot[(3 + u) + (3 + v) * 7], i.e. it
       correlates the image with the slot's 49 floats read as a row-major
       7 x 7 array (row = y offset v, column = x offset u), the same
       convention as our filters F[o][a][b] (a = y offset, b = x offset);
    3. oriMap {4,5,6,7,0,1,2,3} (damascene/src/damascene.cu:598-603): block i
       is copied from slot oriMap[i].

    Returns (block -> our orientation, slot -> our orientation); each slot is
    matched to the one filter of ours equal to it up to sign (sPb sums
    absolute responses), to FP32 rounding."""
    mine = orc.spb_filters().reshape(8, 7, 7)
    slot_ori = {}
    for k in range(8):
        slot = 7 - k
        applied = orc.ref_spb_filter(np.pi / 2 + k * np.pi / 8).astype(np.float64).reshape(7, 7)
        tol = 1e-6 * np.abs(applied).max()
        hits = [o for o in range(8)
                if min(np.abs(mine[o] - applied).max(), np.abs(mine[o] + applied).max()) <= tol]
        assert len(hits) == 1, (k, hits)
        slot_ori[slot] = hits[0]
    ori_map = [4, 5, 6, 7, 0, 1, 2, 3]
    return [slot_ori[ori_map[i]] for i in range(8)], slot_ori


def test_gpb_all_block_order_vs_damascene(orc, gpb):
    """hostGPbAllConcat block i holds our orientation (7 - i) mod 8 (derived
    from the reference's filters, slot rule, kernel indexing and oriMap), and
    gpb_kernel / the oracle's output transform write that order: through the
    oracle's transform, block i of a map whose orientation o is the constant
    o reads (7 - i) mod 8 after the calibration sigmoid is inverted."""
    if not orc.REF_FILTERS.exists():
        pytest.skip("reference filters not built (oracle/_ref needs the reference tree)")
    blocks, slots = damascene_block_orientations(orc)
    assert slots == {s: (3 - s) % 8 for s in range(8)}
    assert blocks == [(7 - i) % 8 for i in range(8)]
    assert blocks == list(gpb.GPB_ALL_BLOCK_ORIENTATION)
    H, W = 3, 5
    g = np.stack([np.full((H, W), 0.01 * (o + 1)) for o in range(8)])
    _, gall = orc.gpb_outputs(g, np.ones((H, W)))
    # the calibration is monotone: rank the blocks by value
    order = np.argsort(gall[:, 0, 0])
    got = [int(np.where(order == i)[0][0]) for i in range(8)]
    assert got == blocks


def test_spb_filters_vs_damascene(orc):
    """The oracle's (and the GPU's) 8 sPb filters are damascene's gaussian_2D
    filters (spectralPb.cu:706-712), which damascene stores x-major at angle
    pi/2 + k pi/8: filter o = +-transpose(damascene k) with o = (8 - k) mod 8
    (the sign is immaterial, sPb sums absolute responses); agreement to FP32
    rounding, the precision damascene builds them in."""
    if not orc.REF_FILTERS.exists():
        pytest.skip("reference filters not built (oracle/_ref needs the reference tree)")
    mine = orc.spb_filters().reshape(8, 7, 7)
    for k in range(8):
        ref = orc.ref_spb_filter(np.pi / 2 + k * np.pi / 8).astype(np.float64).T
        o = (8 - k) % 8
        d = min(np.abs(mine[o] - ref).max(), np.abs(mine[o] + ref).max())
        assert d <= 1e-7 * np.abs(ref).max() * 10, (k, o, d)
        # and no other orientation matches (the mapping is pinned, not just the set)
        for o2 in range(8):
            if o2 != o and not np.allclose(np.abs(mine[o2]), np.abs(mine[o])):
                assert min(np.abs(mine[o2] - ref).max(), np.abs(mine[o2] + ref).max()) > 1e-3

"""Stage (a): Schmid bank, FP32 responses and texton labels.

Oracle pins (CPU): the bank against its closed form (SPFeatures.cpp:117-165),
the declared-order FP32 responses against a direct FP64 correlation with the
REFLECT border (SPFeatures.cpp:285,299-305), the grey conversion against
OpenCV's published BGR2GRAY values, and labels against a brute-force FP64
argmin and a planted-codeword known answer.  GPU: labels bit-exact against the
oracle on the same inputs.
"""
import numpy as np
import pytest

from conftest import synth_packed


def test_grey_known_answers(orc):
    # OpenCV cvtColor(BGR2GRAY) of pure red, green, blue, white, black
    assert orc.lib.orc_grey_u8(0x0000FF) == 76
    assert orc.lib.orc_grey_u8(0x00FF00) == 150
    assert orc.lib.orc_grey_u8(0xFF0000) == 29
    assert orc.lib.orc_grey_u8(0xFFFFFF) == 255
    assert orc.lib.orc_grey_u8(0x000000) == 0


def test_schmid_bank_closed_form(orc):
    bank = orc.schmid_bank()
    x = np.arange(-24, 25, dtype=np.float64)
    X, Y = np.meshgrid(x, x)
    R = np.hypot(X, Y)
    params = [(2, 1), (4, 1), (4, 2), (6, 1), (6, 2), (6, 3), (8, 1), (8, 2), (8, 3), (10, 1), (10, 2), (10, 3), (10, 4)]
    for f, (s, t) in enumerate(params):
        raw = np.cos(R * (3.1415926535 * t / s)) * np.exp(-R * R / (2 * s * s))
        ref = (raw - raw.mean()) / np.abs(raw).sum()
        np.testing.assert_allclose(bank[f], ref, rtol=1e-12, atol=1e-15)
        assert abs(bank[f].sum()) < 1e-12  # zero mean
        assert np.allclose(bank[f], bank[f].T) and np.allclose(bank[f], bank[f][::-1])  # radial symmetry


def test_responses_declared_order_vs_direct(gpb, orc):
    rgb, packed = synth_packed(gpb, 40, 36, regions=6, seed=3)
    cb = np.zeros((13, 32), np.float32)
    _, resp = orc.texton(packed, cb, with_responses=True)
    rng = np.random.default_rng(0)
    for (y, x) in [(0, 0), (35, 39), (17, 20), (0, 39)] + [tuple(p) for p in rng.integers(0, 36, (6, 2))]:
        x = min(x, 39)
        d = orc.schmid_direct(packed, x, y)
        np.testing.assert_allclose(resp[:, y, x], d, rtol=0, atol=2e-3)


def test_labels_vs_bruteforce_argmin(gpb, orc):
    rgb, packed = synth_packed(gpb, 48, 40, regions=8, seed=5)
    cb = gpb.make_codebook(packed, seed=5, iters=10)
    lab, resp = orc.texton(packed, cb, with_responses=True)
    r = resp.reshape(13, -1).T.astype(np.float64)
    d = ((r[:, :, None] - cb[None].astype(np.float64)) ** 2).sum(1)
    bf = d.argmin(1)
    mism = np.nonzero(bf != lab.reshape(-1))[0]
    for i in mism:  # only near-ties may differ between FP32 fmaf and FP64
        dd = np.sort(d[i])
        assert dd[1] - dd[0] <= 1e-5 * max(1.0, dd[0])
    assert len(mism) <= 2


def test_planted_codewords(gpb, orc):
    rgb, packed = synth_packed(gpb, 32, 32, regions=5, seed=7)
    cb0 = np.zeros((13, 32), np.float32)
    _, resp = orc.texton(packed, cb0, with_responses=True)
    pix = [(3, 4), (10, 20), (25, 7), (31, 31)]
    cb = np.full((13, 32), 1e4, np.float32)
    for k, (y, x) in enumerate(pix):
        cb[:, 5 * k + 1] = resp[:, y, x]
    lab = orc.texton(packed, cb)
    for k, (y, x) in enumerate(pix):
        assert lab[y, x] == 5 * k + 1


def test_codebook_deterministic(gpb):
    rgb, packed = synth_packed(gpb, 64, 48, regions=10, seed=11)
    a = gpb.make_codebook(packed, seed=2, iters=10)
    b = gpb.make_codebook(packed, seed=2, iters=10)
    c = gpb.make_codebook(packed, seed=3, iters=10)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.all(np.isfinite(a))


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,seed", [(321, 481, 1), (17, 23, 2), (300, 5, 3), (1, 1, 4), (64, 64, 5)])
def test_texton_gpu_bit_exact(gpb, orc, ctx, w, h, seed):
    print(f"texton parity {w}x{h} seed {seed}", flush=True)
    rgb, packed = synth_packed(gpb, w, h, regions=min(60, w * h), seed=seed)
    cb = gpb.make_codebook(packed, seed=seed, iters=10) if w * h >= 32 else np.random.default_rng(seed).normal(
        0, 20, (13, 32)).astype(np.float32)
    ref = orc.texton(packed, cb)
    got = ctx.texton(packed, cb)
    assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} labels differ"

"""Stage (b): Lab quantisation, half-disc chi-squared cues, mPb, nonmax.

Oracle pins (CPU): Lab bins at known colours and quantiser edge cases,
lib_image::gaussian closed form, chi-squared closed forms (constant labels -> 0;
an ideal step edge gives exactly 8 across it and 0 along it, with and without
bin smoothing), and nonmax on a synthetic ridge.  GPU: every stage against the
oracle on the same inputs — Lab bins exhaustively over all 2^24 colours
(bit-exact), cues and mPb within 1e-4 absolute, nonmax bit-exact given the
same FP32 mPb.
"""
import numpy as np
import pytest

from conftest import synth_packed


def _pack(r, g, b):
    return np.array([[r | (g << 8) | (b << 16)]], np.uint32)


def test_lab_known_answers(orc):
    bins, vals = orc.lab_quantize(_pack(255, 255, 255), with_values=True)
    # white: L = 100 -> 1.0 -> bin clamped from 25 to 24; a, b ~ 0 -> (0+73)/168
    assert bins[0, 0, 0] == 24
    assert abs(vals[0, 0, 0] - 1.0) < 1e-6
    assert abs(vals[1, 0, 0] - 73 / 168) < 1e-3 and abs(vals[2, 0, 0] - 73 / 168) < 1e-3
    bins, vals = orc.lab_quantize(_pack(0, 0, 0), with_values=True)
    assert bins[0, 0, 0] == 0 and vals[0, 0, 0] == 0.0
    # pure red after gamma 2.5 is still (1,0,0): L = 116*0.212671^(1/3)-16 = 53.24
    _, vals = orc.lab_quantize(_pack(255, 0, 0), with_values=True)
    assert abs(vals[0, 0, 0] * 100 - (116 * 0.212671 ** (1 / 3) - 16)) < 1e-9


def test_gaussian_kernel_closed_form(orc):
    for sigma in (2.5, 1.25):
        k = orc.gaussian(sigma)
        s = int(np.ceil(3 * sigma))
        x = np.arange(-s, s + 1)
        ref = np.exp(-x * x / (2 * sigma * sigma))
        np.testing.assert_allclose(k, ref / ref.sum(), rtol=1e-14)


def _step_image(H, W, col, a=3, b=7):
    lab = np.full((H, W), a, np.uint8)
    lab[:, col:] = b
    return lab


@pytest.mark.parametrize("smooth", [None, 2.5])
def test_chi2_step_edge_closed_form(orc, smooth):
    r, B = 5, 6
    lab = _step_image(2 * B + 3, 2 * B + 3, B + 1)
    k = orc.gaussian(smooth) if smooth else None
    out = orc.hist_gradient(lab, B, r, k)  # 3x3 interior
    centre = out[:, 1, 1]  # pixel at column B+1: the edge lies just left of it
    nb = 8
    if k is None:
        expect0 = 8.0
    else:
        # every slice is a single smoothed delta: left = 8 g_b, right = 8 g_a
        s = len(k) // 2
        G = np.zeros((nb, nb))
        for i in range(nb):
            for j in range(nb):
                if abs(i - j) <= s:
                    G[i, j] = k[s + j - i]
        ga, gb = G[:, 3] / G[:, 3].sum(), G[:, 7] / G[:, 7].sum()
        m = (ga - gb) != 0
        expect0 = 0.5 * (64 * (gb - ga)[m] ** 2 / (8 * (ga + gb)[m])).sum()
    assert abs(centre[0] - expect0) < 1e-12
    assert abs(centre[4]) < 1e-12  # split along the edge normal: halves identical


def test_chi2_constant_is_zero(orc):
    lab = np.full((30, 30), 4, np.uint8)
    out = orc.hist_gradient(lab, 10, 10, orc.gaussian(1.25))
    assert np.abs(out).max() < 1e-20  # zero up to the rounding of the spin updates


def test_cues_range_and_orientation(gpb, orc):
    rgb, packed = synth_packed(gpb, 48, 40, regions=6, seed=9)
    lab = orc.texton(packed, gpb.make_codebook(packed, seed=9))
    cues, mpb = orc.local_cues(packed, lab)
    assert cues.min() >= 0 and cues.max() <= 8 + 1e-9
    w = np.array([0.0146, 0.0145, 0.0163, 0.0210, 0.0243, 0.0287, 0.0166, 0.0185, 0.0204, 0.0101, 0.0111, 0.0141])
    np.testing.assert_allclose(mpb, np.tensordot(w, cues, 1), rtol=1e-12, atol=1e-15)


def test_nonmax_ridge(orc):
    # a vertical ridge at column 5 with max orientation 0 survives, flanks vanish
    H, W = 9, 11
    m = np.zeros((8, H, W))
    prof = np.array([0, 0.1, 0.2, 0.3, 0.4, 0.6, 0.4, 0.3, 0.2, 0.1, 0])
    m[0] = prof[None, :]
    nm = orc.nonmax(m)
    assert np.all(nm[1:-1, 5] == 0.6)
    assert np.all(nm[:, :5] == 0) and np.all(nm[:, 6:] == 0)
    assert np.all(nm[0] == 0) and np.all(nm[-1] == 0)  # boundary rows have no interpolation pair
    m2 = m * 2.0  # 1.2*v > 1 clips to 1
    assert np.all(orc.nonmax(m2)[1:-1, 5] == 1.0)


@pytest.mark.gpu
def test_lab_quantize_exhaustive(gpb, orc, ctx):
    print("lab quantize: all 2^24 colours", flush=True)
    allc = np.arange(1 << 24, dtype=np.uint32).reshape(4096, 4096)
    ref = orc.lab_quantize(allc)
    got, nb = ctx.lab_quantize(allc)
    bad = np.count_nonzero(got != ref)
    assert bad == 0, f"{bad} bins differ"
    assert list(nb) == [int(ref[c].max()) + 1 for c in range(3)]


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,seed", [(96, 80, 1), (37, 29, 2), (8, 50, 3)])
def test_cues_gpu_vs_oracle(gpb, orc, ctx, w, h, seed):
    print(f"cues parity {w}x{h} seed {seed}", flush=True)
    rgb, packed = synth_packed(gpb, w, h, regions=8, seed=seed)
    cb = gpb.make_codebook(packed, seed=seed)
    lab = orc.texton(packed, cb)
    cues_ref, mpb_ref = orc.local_cues(packed, lab)
    cues, mpb = ctx.local_cues(packed, lab)
    err_c = np.abs(cues.astype(np.float64) - cues_ref).max()
    err_m = np.abs(mpb.astype(np.float64) - mpb_ref).max()
    print(f"  max |cue err| {err_c:.3e}  max |mPb err| {err_m:.3e}", flush=True)
    assert err_m <= 1e-4
    assert err_c <= 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,seed", [(96, 80, 4), (5, 7, 5)])
def test_nonmax_gpu_bit_exact(gpb, orc, ctx, w, h, seed):
    print(f"nonmax parity {w}x{h} seed {seed}", flush=True)
    rng = np.random.default_rng(seed)
    m = rng.random((8, h, w)).astype(np.float32) * 0.9
    m[:, ::3, :] *= 0.2  # plateaus and ridges
    ref = orc.nonmax(m.astype(np.float64))
    got = ctx.nonmax(m)
    assert np.array_equal(got, ref.astype(np.float32))

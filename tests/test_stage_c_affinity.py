"""Stage (c): intervening-contour affinity W (buildW) and its degree D.

Oracle pins (CPU): closed forms (mPb = 0 gives the same affinity for every
connection of an interior pixel and W_pp = 1; a full-height ridge of mPb = 1
separates the two sides with the affinity of maxpb = 1), W = W^T, D = row
sums, entries in (0, 1]; and a cross-check of the literal restatement of
ic_walk / computeAffinities2 (oracle) against the product's symbolic walk
programs evaluated in numpy, on random mPb including every border case.
GPU: W and D against the oracle on the same mPb, 1e-6 relative.
"""
import math

import numpy as np
import pytest

UP = [(v, u) for u in range(0, 6) for v in range(-5, 6) if u * u + v * v <= 25 and (u > 0 or v > 0)]


def _c_ic_ss(ic):
    val = np.float32(1.8682)
    val = np.float32(np.float64(val) + (np.float64(np.float32(ic)) / 0.3130) * -1.3113)
    post = np.float32(1.0 / (1.0 + math.exp(-np.float64(val))))
    return min(max(post, np.float32(0)), np.float32(1))


def _w_of(maxpb, found):
    sim = np.float32(1.0) - np.float32(maxpb)
    icsim = sim if found else np.float32(0.0)
    pss = _c_ic_ss(np.float32(1.0) - icsim)
    e = np.float32(-(np.float32(1.0) - pss) / np.float32(0.1))
    return np.float32(math.exp(np.float64(e)))


def test_zero_mpb_closed_form(orc):
    H, W = 16, 18
    Wd, D = orc.build_w(np.zeros((H, W)))
    w0 = float(_w_of(0.0, True))
    A = orc.dense_w(Wd)
    assert np.allclose(np.diag(A), 1.0)
    for y in range(5, H - 5):
        for x in range(5, W - 5):
            vals = Wd[:, y, x]
            assert np.allclose(np.delete(vals, 40), w0, rtol=0, atol=1e-7)
    np.testing.assert_array_equal(A, A.T)
    np.testing.assert_allclose(A.sum(1), D.reshape(-1), rtol=1e-15)
    assert A.max() <= 1.0 and A[A > 0].min() > 0


def test_ridge_separates(orc):
    H, W, c = 14, 20, 10
    m = np.zeros((H, W))
    m[:, c] = 1.0
    Wd, _ = orc.build_w(m)
    w_cross, w_same = float(_w_of(1.0, True)), float(_w_of(0.0, True))
    y = 7
    for k, (dx, dy) in enumerate(orc.DISC):
        for x in (c - 3, c + 3):
            x2 = x + dx
            if dx == 0 and dy == 0 or not (0 <= x2 < W) or not (0 <= y + dy < H):
                continue
            crosses = (x < c <= x2) or (x2 < c <= x) or x2 == c or x == c
            if crosses and x2 != c:
                assert abs(Wd[k, y, x] - w_cross) < 1e-7, (k, x)
            elif not crosses:
                assert abs(Wd[k, y, x] - w_same) < 1e-7, (k, x)


def _w_from_programs(gpb, m):
    """Directed affinities from the product's walk programs, symmetrised."""
    H, W = m.shape
    disc = [(v, u) for u in range(-5, 6) for v in range(-5, 6) if u * u + v * v <= 25]
    progs = {}
    A = np.zeros((81, H, W))
    m32 = m.astype(np.float32)
    for y in range(H):
        for x in range(W):
            cfg = (min(x, 5), min(W - 1 - x, 5), min(y, 5), min(H - 1 - y, 5))
            if cfg not in progs:
                progs[cfg] = gpb.ic_program(*cfg)
            valid, found, cnt, start, sites = progs[cfg]
            for k, (dx, dy) in enumerate(disc):
                if k == 40 or not valid[k]:
                    continue
                mx = np.float32(0)
                for s in range(cnt[k]):
                    sdx, sdy = sites[start[k] + s]
                    mx = max(mx, m32[y + sdy, x + sdx])
                A[k, y, x] = _w_of(mx, bool(found[k]))
            A[40, y, x] = 1.0
    Ws = np.zeros_like(A)
    for k, (dx, dy) in enumerate(disc):
        for y in range(H):
            for x in range(W):
                if 0 <= x + dx < W and 0 <= y + dy < H:
                    Ws[k, y, x] = 0.5 * (A[k, y, x] + A[80 - k, y + dy, x + dx])
    return Ws


@pytest.mark.parametrize("H,W,seed", [(14, 17, 0), (7, 9, 1), (3, 12, 2)])
def test_walk_programs_match_literal_walks(gpb, orc, H, W, seed):
    rng = np.random.default_rng(seed)
    m = rng.random((H, W)) * (rng.random((H, W)) < 0.3)
    Wd, D = orc.build_w(m)
    Ws = _w_from_programs(gpb, m)
    np.testing.assert_allclose(Ws, Wd, rtol=1e-6, atol=1e-12)


def _upper(Wd):
    idx = [k for k, (dx, dy) in enumerate([(v, u) for u in range(-5, 6) for v in range(-5, 6) if u * u + v * v <= 25])
           if dy > 0 or (dy == 0 and dx > 0)]
    return Wd[idx]


@pytest.mark.gpu
@pytest.mark.parametrize("H,W,seed", [(48, 64, 0), (7, 9, 1), (30, 3, 2), (101, 37, 3)])
def test_build_w_gpu(gpb, orc, ctx, H, W, seed):
    print(f"build_w parity {W}x{H} seed {seed}", flush=True)
    rng = np.random.default_rng(seed)
    m = (rng.random((H, W)) * (rng.random((H, W)) < 0.3)).astype(np.float32)
    Wd, D = orc.build_w(m.astype(np.float64))
    wu, Dg = ctx.build_w(m)
    np.testing.assert_allclose(wu, _upper(Wd), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(Dg, D, rtol=1e-6)

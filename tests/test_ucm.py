"""OWT-UCM, the consumer of the 8 oriented gPb maps (src/contours2ucm.cpp).

The oracle (oracle/ucm_oracle.py) restates the reference's steps; its contour
step is the reference's own lib_image contour_set (oracle/_ref).  Its other
steps are pinned here by what fixes them independently:
  * watershed: the invariants of the level-by-level raster flooding (no two
    adjacent pixels with different basins, every watershed pixel between two
    basins, connected basins) and a two-valley image whose watershed line is
    known;
  * 8-bit scaling: the end points and round-half-to-even;
  * chord orientation: horizontal, vertical and diagonal chords;
  * doubled grid: a hand-computed 2 x 3 case;
  * bwmorph clean and bwlabel: scipy's neighbour counts and labelling;
  * cleaning: a hand-computed isolated zero;
  * merging: an independent greedy merge that recomputes every mean boundary
    from scratch (each boundary element ends at the strength of the merge
    that joined its two sides), and the corner-site completion;
  * calibration: its values at 0 and 1 and its monotonicity.
The product (host C++ behind gpb_contours2ucm and friends) is then compared
with the oracle bit for bit, stage by stage; the contour step also directly
against the reference's compiled code on many binary images.  The GPU test
runs the product on gPb maps the GPU path computed.
"""
import importlib.util
import math
import sys

import numpy as np
import pytest
from scipy import ndimage

from conftest import ROOT, synth_packed


@pytest.fixture(scope="module")
def uo():
    path = ROOT / "oracle" / "ucm_oracle.py"
    if not (ROOT / "oracle" / "_ref" / "libref_libimage.so").exists():
        pytest.skip("reference lib_image not built (oracle/_ref needs the reference tree)")
    name = "ucm_oracle_mod"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, path)
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    return sys.modules[name]


def _smooth_maps(seed, H, W):
    rng = np.random.default_rng(seed)
    s = rng.uniform(1, 3)
    pbo = np.stack([np.abs(ndimage.gaussian_filter(rng.standard_normal((H, W)), s)) for _ in range(8)])
    return pbo / pbo.max()


def _smooth_u8(seed, H, W, sigma):
    rng = np.random.default_rng(seed)
    f = ndimage.gaussian_filter(rng.standard_normal((H, W)), sigma)
    return np.clip((f - f.min()) / (f.max() - f.min()) * 255, 0, 255).round().astype(np.uint8)


# ---- oracle pins -------------------------------------------------------------------


def test_watershed_invariants(uo):
    for seed in range(6):
        img = _smooth_u8(seed, 40 + 3 * seed, 35, 1.5 + 0.3 * seed)
        L = uo.watershed_yupeng(img)
        H, W = L.shape
        assert L.min() >= 0
        pad = np.pad(L, 1, constant_values=0)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                if dx == dy == 0:
                    continue
                nb = pad[1 + dx:1 + dx + H, 1 + dy:1 + dy + W]
                # two adjacent basin pixels always share their basin
                assert not np.any((L > 0) & (nb > 0) & (nb != L))
        # every watershed pixel separates at least two basins
        for x, y in np.argwhere(L == 0):
            nb = {int(v) for v in L[max(0, x - 1):x + 2, max(0, y - 1):y + 2].ravel() if v > 0}
            assert len(nb) >= 2
        # basins are connected
        for b in np.unique(L[L > 0]):
            _, n = ndimage.label(L == b, structure=np.ones((3, 3)))
            assert n == 1


def test_watershed_two_valleys(uo):
    # two V-shaped valleys along the columns, a ridge at column 7: the
    # watershed line is that column
    col = np.array([5, 4, 3, 2, 1, 2, 3, 9, 3, 2, 1, 2, 3, 4, 5], np.uint8)
    img = np.tile(col, (6, 1))
    L = uo.watershed_yupeng(img)
    assert np.all(L[:, 7] == 0)
    assert len(np.unique(L[:, :7])) == 1 and len(np.unique(L[:, 8:])) == 1 and L[0, 0] != L[0, 14]


def test_pb_to_u8(uo):
    pb = np.array([[0.0, 2.5, 3.5], [4.5, 100.0, 255.0]])
    np.testing.assert_array_equal(uo.pb_to_u8(pb), [[0, 2, 4], [4, 100, 255]])
    pb2 = np.array([[1.0, 1.5], [2.0, 3.0]])  # lo -> 0, hi -> 255
    q = uo.pb_to_u8(pb2)
    assert q[0, 0] == 0 and q[1, 1] == 255 and q[1, 0] == round(255 / 2)


def test_chord_orientation(uo):
    # (row, col) end points: a horizontal chord (angle 0) is orientation 4, a
    # vertical one 0, rising diagonals 2 (45 deg) and falling 6 (-45 deg)
    assert uo._orient((5, 1), (5, 9)) == 4
    assert uo._orient((1, 4), (9, 4)) == 0
    assert uo._orient((1, 1), (4, 4)) == 2
    assert uo._orient((1, 4), (4, 1)) == 6
    assert uo._orient((0, 0), (1, 3)) == 3   # atan(1/3) = 18.4 deg
    assert uo._orient((0, 0), (3, 1)) == 1   # atan(3) = 71.6 deg


def test_super_contour_4c(uo):
    pb = np.array([[1.0, 2.0, 3.0], [4.0, 0.0, 6.0]])
    # pixels at even sites, the min of horizontal / vertical neighbours
    # between them, last row copied, last column the max with its neighbour
    want = np.array([
        [1, 1, 2, 2, 3, 3],
        [1, 0, 0, 0, 3, 3],
        [4, 0, 0, 0, 6, 6],
        [4, 0, 0, 0, 6, 6],
    ], float)
    np.testing.assert_array_equal(uo.super_contour_4c(pb), want)


def test_bwmorph_clean_and_bwlabel(uo):
    rng = np.random.default_rng(3)
    b = (rng.random((30, 27)) < 0.15).astype(np.uint8)
    cnt = ndimage.convolve(b.astype(int), np.ones((3, 3), int), mode="constant") - b
    np.testing.assert_array_equal(uo.bwmorph_clean(b), b * (cnt > 0))
    lab, n = ndimage.label(b, structure=np.ones((3, 3)))
    first = {}
    for v in lab.ravel():
        if v and v not in first:
            first[v] = len(first) + 1
    want = np.vectorize(lambda v: first.get(v, 0))(lab)
    np.testing.assert_array_equal(uo.bwlabel(b), want)


def test_clean_watersheds_isolated_zero(uo):
    # a zero at (3, 3) with no zero among its 8 neighbours: of its four
    # diagonal directions the weakest is down-right (max(ws[5,4], ws[4,5]) =
    # 2); as ws[5,4] < ws[4,5], ws[3,4] becomes 0 and ws[4,3] and ws[4,4] 2
    ws = np.full((8, 8), 9.0)
    ws[3, 3] = 0
    ws[5, 4], ws[4, 5] = 1.0, 2.0
    out = uo.clean_watersheds(ws)
    want = ws.copy()
    want[3, 4], want[4, 3], want[4, 4] = 0.0, 2.0, 2.0
    np.testing.assert_array_equal(out, want)


def _naive_ucm_levels(local, part):
    """Independent greedy merging: the adjacent pair of current regions with
    the smallest mean boundary strength (ties: smaller ids) merges, means
    recomputed from all boundary elements at every step.  Returns, per
    boundary element (column-major doubled-grid index), the strength of the
    merge that joined its two sides."""
    tx, ty = part.shape
    txc = 2 * tx + 1
    elems = {}
    for x in range(tx):
        for y in range(ty):
            for dx, dy in ((1, 0), (0, 1)):
                nx, ny = x + dx, y + dy
                if nx < tx and ny < ty and part[x, y] != part[nx, ny]:
                    elems[(x + nx + 1) + (y + ny + 1) * txc] = (int(part[x, y]), int(part[nx, ny]))
    rep = {int(r): int(r) for r in np.unique(part)}
    find = lambda r: r if rep[r] == r else find(rep[r])
    level = {}
    while True:
        pairs = {}
        for coord, (a, b) in elems.items():
            ra, rb = find(a), find(b)
            if ra != rb:
                k = (min(ra, rb), max(ra, rb))
                pairs.setdefault(k, []).append(local[coord])
        if not pairs:
            return level
        (ra, rb), vals = min(pairs.items(), key=lambda kv: (sum(kv[1]) / len(kv[1]), kv[0]))
        m = sum(vals) / len(vals)
        for coord, (a, b) in elems.items():
            if {find(a), find(b)} == {ra, rb}:
                level[coord] = m
        rep[min(ra, rb)] = max(ra, rb)


def test_ucm_mean_pb_against_naive_merging(uo):
    rng = np.random.default_rng(5)
    for trial in range(3):
        tx, ty = 9, 11
        # a small partition: the 4-connected pieces of a blocky random field
        field = ndimage.zoom(rng.integers(0, 5, (3, 4)), (3, 3), order=0)[:tx, :ty]
        parts = np.zeros((tx, ty), np.int64)
        n = 0
        for v in np.unique(field):
            lab, k = ndimage.label(field == v, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
            parts[lab > 0] = lab[lab > 0] - 1 + n
            n += k
        ws_wt2 = rng.random((2 * tx + 1, 2 * ty + 1))
        ucm = uo.ucm_mean_pb(ws_wt2, parts)
        levels = _naive_ucm_levels(ws_wt2.reshape(-1, order="F"), parts)
        flat = ucm.reshape(-1, order="F")
        assert n > 2 and len(levels) > 0
        for coord, m in levels.items():
            assert abs(flat[coord] - m) <= 1e-12
        # pixel sites stay 0; a corner site is the max of its four boundary elements
        assert np.all(ucm[1::2, 1::2] == 0)
        assert ucm[2, 2] == max(ucm[1, 2], ucm[3, 2], ucm[2, 1], ucm[2, 3])


def test_normalize_output(uo):
    v = np.array([[0.0, 0.1, 0.5, 1.0]])
    out = uo.normalize_output(v)
    assert out[0, 0] == 0.0  # 1 / (1 + e^2.7487) = 0.06015 < 0.0602
    p1 = 1 / (1 + math.exp(-(11.1189 - 2.7487)))
    assert abs(out[0, 3] - (p1 - 0.0602) / (1 - 0.0602)) <= 1e-15
    assert np.all(np.diff(out[0]) >= 0) and out.max() <= 1


# ---- product against oracle and reference ------------------------------------------


def test_watershed_product_vs_oracle(gpb, uo):
    for seed in range(4):
        img = _smooth_u8(seed, 37, 52, 1.0 + 0.5 * seed)
        np.testing.assert_array_equal(gpb.watershed8(img), uo.watershed_yupeng(img))


def test_fit_contours_product_vs_reference(gpb, uo):
    rng = np.random.default_rng(11)
    for trial in range(60):
        H, W = rng.integers(20, 110, 2)
        kind = trial % 3
        if kind == 0:
            bw = uo.watershed_yupeng(_smooth_u8(trial, H, W, rng.uniform(0.8, 4))) == 0
        elif kind == 1:
            bw = rng.random((H, W)) < rng.uniform(0.05, 0.5)
        else:
            f = ndimage.gaussian_filter(rng.standard_normal((H, W)), 2.0)
            bw = np.abs(f) < rng.uniform(0.01, 0.05)
        mine, ref = gpb.fit_contours(bw), uo.ref_contours(bw)
        assert (mine is None) == (ref is None), trial
        if ref is not None:
            for a, b in zip(mine, ref):
                np.testing.assert_array_equal(a, b, err_msg=f"trial {trial}")


def test_contours2ucm_product_vs_oracle(gpb, uo):
    for seed in range(5):
        H, W = 30 + 5 * seed, 44 - 3 * seed
        pbo = _smooth_maps(seed, H, W)
        st = gpb.ucm_stages(pbo)
        ucm, ref = uo.contours2ucm(pbo, True)
        for k in ("pb8", "ws", "ws_wt", "ws_wt2", "labels", "super_ucm"):
            np.testing.assert_array_equal(st[k], ref[k], err_msg=f"seed {seed} stage {k}")
        assert st["regions"] == ref["regions"] > 1
        np.testing.assert_array_equal(gpb.contours2ucm(pbo, True), ucm)
        np.testing.assert_array_equal(gpb.contours2ucm(pbo, False), uo.contours2ucm(pbo, False)[0])


@pytest.mark.gpu
def test_ucm_on_gpu_gpb_gpu(gpb, uo, ctx):
    """The consumer on gPb maps computed by the GPU path (a 96 x 80 Voronoi
    image): product and oracle agree bit for bit on every stage."""
    _, packed = synth_packed(gpb, 96, 80, regions=14, sigma=6.0, seed=7)
    cb = gpb.make_codebook(packed, seed=7, iters=10)
    ctx.set_eig_options(nvec=9, basis=40, keep=20, tol=1e-6)
    out = ctx.compute(packed, cb)
    ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=1e-3)
    pbo = out["gpb_orient"].astype(np.float64)
    st = gpb.ucm_stages(pbo)
    ucm, ref = uo.contours2ucm(pbo, True)
    print(f"96x80 gPb -> {st['regions']} initial regions", flush=True)
    for k in ("pb8", "ws", "ws_wt", "ws_wt2", "labels", "super_ucm"):
        np.testing.assert_array_equal(st[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(gpb.contours2ucm(pbo, True), ucm)


class _Hip:
    """Device buffers through the HIP runtime the library itself uses."""

    def __init__(self):
        import ctypes
        self.C = ctypes
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.C.c_void_p()
        assert self.lib.hipMalloc(self.C.byref(p), a.nbytes) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def free(self, p):
        assert self.lib.hipFree(p) == 0


@pytest.mark.gpu
def test_ucm_device_paths_gpu(gpb, uo, ctx):
    """gpb_contours2ucm_device (the per-pixel stages on the GPU) and
    gpb_compute_ucm (gPb then UCM, all on one device image) against the host
    consumer on the same maps, both output sizes; the host consumer against
    the oracle on them."""
    W, H = 120, 90
    _, packed = synth_packed(gpb, W, H, regions=18, sigma=6.0, seed=9)
    cb = gpb.make_codebook(packed, seed=9, iters=10)
    out = ctx.compute(packed, cb)
    pbo = out["gpb_orient"]
    hip = _Hip()
    d_pbo = hip.upload(pbo.astype(np.float32))
    d_rgb = hip.upload(packed.astype(np.uint32))
    d_cb = hip.upload(cb.astype(np.float32))
    try:
        for dbl in (True, False):
            host = gpb.contours2ucm(pbo.astype(np.float64), dbl)
            print(f"{W}x{H}: contours2ucm_device, double_size={dbl}", flush=True)
            np.testing.assert_array_equal(ctx.contours2ucm_device(d_pbo.value, W, H, dbl), host)
            print(f"{W}x{H}: compute_ucm, double_size={dbl}", flush=True)
            np.testing.assert_array_equal(ctx.compute_ucm(W, H, d_rgb.value, d_cb.value, dbl), host)
        np.testing.assert_array_equal(gpb.contours2ucm(pbo.astype(np.float64), True),
                                      uo.contours2ucm(pbo.astype(np.float64), True)[0])
    finally:
        for p in (d_pbo, d_rgb, d_cb):
            hip.free(p)


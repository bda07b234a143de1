"""The reference-named C++ entry points (include/gpb_ucm.hpp) and gpb_cli.

CPU: libgpb.so exports the seven reference names with the reference's
argument lists (computeGPb / getCudaDeviceCount with C++ linkage in the
global namespace, as damascene/include/globalPb.h:37,44 declares them); the
OpenCV-YAML writer's output parses back to the matrix written; a bad image
is rejected with exit status 2 before any GPU work.  GPU: every entry point,
driven through gpb_cli, reproduces the C-ABI results (which the stage tests
pin against the oracle) on the same image: cues, textons, nonmax and the
buildW triplets bit-exact, sPb / gPb to 1e-4 (the eigensolves run in another
process, whose SpMV tile may be tuned differently, so the FP64 reductions
may round differently).
"""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import synth_packed

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "ucm-mpi-gpu_amd"
CLI = PKG / "gpb_cli"


def _cli(*args, check=True):
    r = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=300)
    if check and r.returncode != 0:
        raise AssertionError(f"gpb_cli {args} -> {r.returncode}\n{r.stdout}\n{r.stderr}")
    return r


def _parse_yaml(path):
    """name -> 2-D array from OpenCV FileStorage YAML (!!opencv-matrix, dt: d)."""
    text = Path(path).read_text()
    assert text.startswith("%YAML:1.0")
    out = {}
    for m in re.finditer(r"(\w+): !!opencv-matrix\s+rows: (\d+)\s+cols: (\d+)\s+dt: d\s+data: \[([^\]]*)\]", text):
        vals = np.array([float(v) for v in m.group(4).replace("\n", " ").split(",")])
        out[m.group(1)] = vals.reshape(int(m.group(2)), int(m.group(3)))
    return out


def _write_ppm(path, rgb):
    h, w, _ = rgb.shape
    Path(path).write_bytes(b"P6\n%d %d\n255\n" % (w, h) + np.ascontiguousarray(rgb, np.uint8).tobytes())


def test_reference_names_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", str(PKG / "libgpb.so")], capture_output=True, text=True,
                         check=True).stdout
    for sig in [
        "computeGPb(unsigned int, unsigned int, unsigned int, unsigned int*, float**, float**)",
        "getCudaDeviceCount()",
        "gpb_ucm::det_mPb(gpb_ucm::detmpb&, gpb_ucm::Mat&)",
        "gpb_ucm::multiscalePb(gpb_ucm::Mat*&, gpb_ucm::Mat*&, gpb_ucm::detmpb&, gpb_ucm::Mat&, gpb_ucm::Size, int, ",
        "gpb_ucm::spectralPb(std::vector<gpb_ucm::Mat*, std::allocator<gpb_ucm::Mat*> >&, gpb_ucm::Mat*, "
        "gpb_ucm::Size, ",
        "gpb_ucm::buildW(std::vector<double, std::allocator<double> >&, std::vector<double, std::allocator<double> >&,"
        " std::vector<int, std::allocator<int> >&, std::vector<int, std::allocator<int> >&, gpb_ucm::Mat&, "
        "gpb_ucm::Mat&)",
        "gpb_ucm::GlobalPb(",
        "gpb_ucm::Contours2Ucm(",
    ]:
        assert sig in out, sig


def test_cli_yaml_format(tmp_path):
    _cli("yaml", tmp_path / "m.yml")
    m = _parse_yaml(tmp_path / "m.yml")["m"]
    np.testing.assert_array_equal(m, (0.5 * np.arange(6) - 1.0 / 3.0).reshape(2, 3))


def test_cli_contours2ucm_matches_capi(gpb, tmp_path):
    """gpb_ucm::Contours2Ucm (through gpb_cli) against gpb_contours2ucm, both
    output sizes."""
    from scipy import ndimage
    rng = np.random.default_rng(4)
    H, W = 33, 41
    pbo = np.stack([np.abs(ndimage.gaussian_filter(rng.standard_normal((H, W)), 1.8)) for _ in range(8)])
    (tmp_path / "maps.f64").write_bytes(np.ascontiguousarray(pbo, np.float64).tobytes())
    for fmt, shape in (("doubleSize", (2 * H + 1, 2 * W + 1)), ("", (H + 1, W + 1))):
        _cli("ucm", tmp_path / "maps.f64", W, H, tmp_path / "u.f64", *([fmt] if fmt else []))
        got = np.fromfile(tmp_path / "u.f64", np.float64).reshape(shape)
        np.testing.assert_array_equal(got, gpb.contours2ucm(pbo, fmt == "doubleSize"))


def test_cli_rejects_bad_image(tmp_path):
    (tmp_path / "bad.ppm").write_bytes(b"P5\n4 4\n255\n" + bytes(16))  # greyscale PGM, not P6
    r = _cli("gpb", tmp_path / "bad.ppm", tmp_path / "out", check=False)
    assert r.returncode == 2 and "PPM" in r.stderr
    r = _cli("compute", tmp_path / "missing.ppm", tmp_path / "o.bin", check=False)
    assert r.returncode == 2


@pytest.mark.gpu
def test_shims_match_capi_gpu(gpb, ctx, tmp_path):
    W, H = 72, 56
    print(f"shims {W}x{H}", flush=True)
    rgb, packed = synth_packed(gpb, W, H, regions=9, seed=11)
    _write_ppm(tmp_path / "img.ppm", rgb)
    _cli("shims", tmp_path / "img.ppm", tmp_path)
    N = W * H
    ld = lambda name, dt=np.float64: np.fromfile(tmp_path / name, dt)

    cb = gpb.make_codebook(packed, seed=1, iters=10, max_samples=65536)
    t = ctx.texton(packed, cb)
    np.testing.assert_array_equal(ld("textons.f64").reshape(H, W), t)
    cues, mpb_all = ctx.local_cues(packed, t)
    np.testing.assert_array_equal(ld("cues.f64").reshape(12, 8, H, W), cues)
    nmax = ctx.nonmax(mpb_all)
    np.testing.assert_array_equal(ld("nmax.f64").reshape(H, W), nmax)

    # buildW triplets: diagonal then the upper offsets in row-major order
    wup, deg = ctx.build_w(nmax)
    offs = [(dy, dx) for dy in range(6) for dx in range(-5, 6) if dx * dx + dy * dy <= 25 and (dy > 0 or dx > 0)]
    I, J, V = [], [], []
    for y in range(H):
        for x in range(W):
            p = y * W + x
            I.append(p), J.append(p), V.append(1.0)
            for k, (dy, dx) in enumerate(offs):
                if 0 <= x + dx < W and y + dy < H:
                    I.append(p), J.append((y + dy) * W + x + dx), V.append(float(wup[k, y, x]))
    np.testing.assert_array_equal(ld("I.i32", np.int32), I)
    np.testing.assert_array_equal(ld("J.i32", np.int32), J)
    np.testing.assert_array_equal(ld("val.f64"), V)
    np.testing.assert_array_equal(ld("valD.f64"), deg.reshape(-1))
    Ia, Ja = ld("I.i32", np.int32), ld("J.i32", np.int32)
    same = Ia[1:] == Ia[:-1]
    assert np.all(np.diff(Ia) >= 0) and np.all(Ja[1:][same] > Ja[:-1][same]) and np.all(Ja >= Ia)

    for nvec in (9, 17):
        keep = max(20, nvec + 8)
        ctx.set_eig_options(nvec=nvec, basis=max(40, 2 * keep), keep=keep, tol=1e-3)
        spb, _ = ctx.spectral_pb(nmax)
        got = ld(f"spb{nvec}.f64").reshape(8, H, W)
        err = np.abs(got - spb).max() / np.abs(spb).max()
        print(f"  spectralPb nvec {nvec}: rel diff {err:.2e}", flush=True)
        assert err <= 1e-4
        if nvec == 17:
            out17 = ctx.compute(packed, cb)
    g = ld("globalpb.f64").reshape(8, H, W)
    assert np.abs(g - out17["gpb_orient"]).max() <= 1e-4
    ymls = _parse_yaml(tmp_path / "img_gPb.yml")
    for o in range(8):
        np.testing.assert_array_equal(ymls[f"gPb_{o}"], g[o])
    np.testing.assert_array_equal(_parse_yaml(tmp_path / "img_mPb.yml")["mPb_rsz_mat"], nmax)

    ctx.set_eig_options(nvec=9, basis=40, keep=20, tol=1e-3)
    out9 = ctx.compute(packed, cb)
    gmax, gall = ld("gmax.f32", np.float32).reshape(H, W), ld("gall.f32", np.float32).reshape(8, W, H)
    d1, d2 = np.abs(gmax - out9["gpb_max"]).max(), np.abs(gall - out9["gpb_all"]).max()
    print(f"  computeGPb: |d gmax| {d1:.2e} |d gall| {d2:.2e}", flush=True)
    assert d1 <= 1e-4 and d2 <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("edge", ["vertical", "horizontal"])
def test_compute_gpb_block_order_gpu(tmp_path, edge):
    """computeGPb (through gpb_cli) on a step edge: at the pixels where the
    masked gPb maximum lies on the edge, the largest hostGPbAllConcat block is
    the one that holds the edge's orientation in damascene's order, block
    7 - o for orientation o (a vertical edge is orientation 0, a horizontal
    one orientation 4: the half-disc split runs along the edge)."""
    W, H = 64, 48
    print(f"computeGPb block order, {edge} edge {W}x{H}", flush=True)
    rng = np.random.default_rng(21)
    yy, xx = np.mgrid[0:H, 0:W]
    step = (xx >= W // 2) if edge == "vertical" else (yy >= H // 2)
    img = np.where(step, 190.0, 70.0)[..., None] + rng.normal(0, 3.0, (H, W, 3))
    rgb = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    _write_ppm(tmp_path / "step.ppm", rgb)
    _cli("compute", tmp_path / "step.ppm", tmp_path / "out.bin")
    raw = np.fromfile(tmp_path / "out.bin", np.float32)
    gmax = raw[:W * H].reshape(H, W)
    gall = raw[W * H:].reshape(8, W, H).transpose(0, 2, 1)  # block, y, x
    o = 0 if edge == "vertical" else 4
    want = 7 - o
    if edge == "vertical":
        ys = np.arange(8, H - 8)
        xs = gmax[ys].argmax(axis=1)
        assert np.all(np.abs(xs - (W // 2 - 0.5)) <= 1.5), xs
    else:
        xs = np.arange(8, W - 8)
        ys = gmax[:, xs].argmax(axis=0)
        assert np.all(np.abs(ys - (H // 2 - 0.5)) <= 1.5), ys
    blocks = gall[:, ys, xs].argmax(axis=0)
    assert np.all(blocks == want), np.bincount(blocks, minlength=8)

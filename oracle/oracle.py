"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


if not LIB_PATH.exists():
    build()
lib = C.CDLL(str(LIB_PATH))

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("orc_grey_u8", C.c_int, [C.c_uint32])
_sig("orc_schmid_bank", None, [_f64p])
_sig("orc_schmid_direct", None, [_u32p, C.c_int, C.c_int, C.c_int, C.c_int, _f64p])
_sig("orc_texton", None, [_u32p, C.c_int, C.c_int, _f32p, _u8p, C.c_void_p])
_sig("orc_lab_quantize", None, [_u32p, C.c_size_t, C.c_int, _u8p, C.c_void_p])
_sig("orc_mirror_u8", None, [_u8p, C.c_int, C.c_int, C.c_int, _u8p])
_sig("orc_gaussian", C.c_int, [C.c_double, _f64p])
_sig("orc_hist_gradient", None, [_u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, _f64p])
_sig("orc_local_cues", None, [_u32p, _u8p, C.c_int, C.c_int, _f64p, C.c_void_p])
_sig("orc_local_cues_nb", None, [_u32p, _u8p, C.c_int, C.c_int, C.c_void_p, _f64p, C.c_void_p])
_sig("orc_nonmax", None, [_f64p, C.c_int, C.c_int, _f64p])


def grey(packed: np.ndarray) -> np.ndarray:
    f = np.vectorize(lambda v: lib.orc_grey_u8(int(v)), otypes=[np.int32])
    return f(packed)


def schmid_bank() -> np.ndarray:
    out = np.empty((13, 49, 49), np.float64)
    lib.orc_schmid_bank(out.reshape(-1))
    return out


def schmid_direct(packed: np.ndarray, x: int, y: int) -> np.ndarray:
    packed = np.ascontiguousarray(packed, np.uint32)
    out = np.empty(13, np.float64)
    lib.orc_schmid_direct(packed.reshape(-1), packed.shape[1], packed.shape[0], x, y, out)
    return out


def texton(packed: np.ndarray, codebook: np.ndarray, with_responses: bool = False):
    packed = np.ascontiguousarray(packed, np.uint32)
    H, W = packed.shape
    lab = np.empty((H, W), np.uint8)
    resp = np.empty((13, H, W), np.float32) if with_responses else None
    lib.orc_texton(packed.reshape(-1), W, H, np.ascontiguousarray(codebook, np.float32).reshape(-1), lab.reshape(-1),
                   resp.ctypes.data if resp is not None else None)
    return (lab, resp) if with_responses else lab


def lab_quantize(packed: np.ndarray, nbins: int = 25, with_values: bool = False):
    packed = np.ascontiguousarray(packed, np.uint32)
    n = packed.size
    bins = np.empty((3,) + packed.shape, np.uint8)
    vals = np.empty((3,) + packed.shape, np.float64) if with_values else None
    lib.orc_lab_quantize(packed.reshape(-1), n, nbins, bins.reshape(-1), vals.ctypes.data if vals is not None else None)
    return (bins, vals) if with_values else bins


def mirror_u8(img: np.ndarray, border: int) -> np.ndarray:
    img = np.ascontiguousarray(img, np.uint8)
    H, W = img.shape
    out = np.empty((H + 2 * border, W + 2 * border), np.uint8)
    lib.orc_mirror_u8(img.reshape(-1), W, H, border, out.reshape(-1))
    return out


def gaussian(sigma: float) -> np.ndarray:
    buf = np.empty(2 * int(np.ceil(3 * sigma)) + 8, np.float64)
    n = lib.orc_gaussian(sigma, buf)
    return buf[:n].copy()


def hist_gradient(labels_mirrored: np.ndarray, border: int, r: int, kernel: np.ndarray | None) -> np.ndarray:
    lm = np.ascontiguousarray(labels_mirrored, np.uint8)
    Hm, Wm = lm.shape
    out = np.empty((8, Hm - 2 * border, Wm - 2 * border), np.float64)
    if kernel is not None:
        k = np.ascontiguousarray(kernel, np.float64)
        lib.orc_hist_gradient(lm.reshape(-1), Wm, Hm, border, r, k.ctypes.data, k.size, out.reshape(-1))
    else:
        lib.orc_hist_gradient(lm.reshape(-1), Wm, Hm, border, r, None, 0, out.reshape(-1))
    return out


def local_cues(packed: np.ndarray, textons: np.ndarray, nbins=None):
    """Returns cues[12, 8, H, W] and mpb_all[8, H, W] (FP64).  nbins (optional,
    4 ints): the histogram lengths of L, a, b and textons to use instead of
    max(label) + 1 over this input (those of a whole image, for a crop)."""
    packed = np.ascontiguousarray(packed, np.uint32)
    H, W = packed.shape
    cues = np.empty((12, 8, H, W), np.float64)
    mpb = np.empty((8, H, W), np.float64)
    t = np.ascontiguousarray(textons, np.uint8).reshape(-1)
    if nbins is None:
        lib.orc_local_cues(packed.reshape(-1), t, W, H, cues.reshape(-1), mpb.ctypes.data)
    else:
        nb = np.ascontiguousarray(nbins, np.int32)
        lib.orc_local_cues_nb(packed.reshape(-1), t, W, H, nb.ctypes.data, cues.reshape(-1), mpb.ctypes.data)
    return cues, mpb


def nonmax(mpb_all: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(mpb_all, np.float64)
    _, H, W = m.shape
    out = np.empty((H, W), np.float64)
    lib.orc_nonmax(m.reshape(-1), W, H, out.reshape(-1))
    return out

_sig("orc_build_w", C.c_int, [_f64p, C.c_int, C.c_int, _f64p, _f64p])
_sig("orc_eigs", C.c_int, [_f64p, _f64p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _f64p, _f64p,
                           C.POINTER(C.c_int)])
_sig("orc_gen_residual", C.c_double, [_f64p, _f64p, C.c_int, C.c_int, C.c_double, _f64p])
_sig("orc_jacobi_eig", None, [C.c_int, _f64p, _f64p, _f64p])
_sig("orc_spb_filters", None, [_f64p])
_sig("orc_spb", None, [_f64p, _f64p, C.c_int, C.c_int, C.c_int, _f64p])
_sig("orc_gpb", None, [_f64p, _f64p, C.c_size_t, _f64p])
_sig("orc_gpb_outputs", None, [_f64p, _f64p, C.c_int, C.c_int, _f32p, _f32p])

DISC = [(v, u) for u in range(-5, 6) for v in range(-5, 6) if u * u + v * v <= 25]  # (dx, dy), row-major


def build_w(mpb_nmax: np.ndarray):
    """Returns (Wd[81, H, W], D[H, W]) for the 81 disc offsets DISC."""
    m = np.ascontiguousarray(mpb_nmax, np.float64)
    H, W = m.shape
    Wd = np.empty((81, H, W), np.float64)
    D = np.empty((H, W), np.float64)
    rc = lib.orc_build_w(m.reshape(-1), W, H, Wd.reshape(-1), D.reshape(-1))
    if rc != 0:
        raise RuntimeError("support scan line not strictly increasing")
    return Wd, D


def dense_w(Wd: np.ndarray) -> np.ndarray:
    """Dense N x N W from the disc layout (tiny images only)."""
    _, H, W = Wd.shape
    N = H * W
    A = np.zeros((N, N))
    for k, (dx, dy) in enumerate(DISC):
        for y in range(H):
            for x in range(W):
                x2, y2 = x + dx, y + dy
                if 0 <= x2 < W and 0 <= y2 < H:
                    A[y * W + x, y2 * W + x2] = Wd[k, y, x]
    return A


def eigs(Wd: np.ndarray, D: np.ndarray, K: int = 9, m: int = 40, tol: float = 1e-10, maxit: int = 100000):
    Wd = np.ascontiguousarray(Wd, np.float64)
    D = np.ascontiguousarray(D, np.float64)
    H, W = D.shape
    ev = np.empty(K, np.float64)
    vec = np.empty((K, H, W), np.float64)
    it = C.c_int(0)
    rc = lib.orc_eigs(Wd.reshape(-1), D.reshape(-1), W, H, K, m, tol, maxit, ev, vec.reshape(-1), C.byref(it))
    if rc < 0:
        raise RuntimeError("oracle eigensolver did not converge")
    return ev, vec, it.value


def gen_residual(Wd, D, lam, v) -> float:
    H, W = D.shape
    return lib.orc_gen_residual(np.ascontiguousarray(Wd, np.float64).reshape(-1),
                                np.ascontiguousarray(D, np.float64).reshape(-1), W, H, float(lam),
                                np.ascontiguousarray(v, np.float64).reshape(-1))


def spb_filters() -> np.ndarray:
    out = np.empty((8, 7, 7), np.float64)
    lib.orc_spb_filters(out.reshape(-1))
    return out


def spb(evecs: np.ndarray, evals: np.ndarray) -> np.ndarray:
    e = np.ascontiguousarray(evecs, np.float64)
    K, H, W = e.shape
    out = np.empty((8, H, W), np.float64)
    lib.orc_spb(e.reshape(-1), np.ascontiguousarray(evals, np.float64), K, W, H, out.reshape(-1))
    return out


def gpb(cues: np.ndarray, spb_maps: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(cues, np.float64)
    s = np.ascontiguousarray(spb_maps, np.float64)
    _, H, W = s.shape
    out = np.empty((8, H, W), np.float64)
    lib.orc_gpb(c.reshape(-1), s.reshape(-1), H * W, out.reshape(-1))
    return out


def gpb_outputs(gpb_maps: np.ndarray, mpb_nmax: np.ndarray):
    """computeGPb outputs: (gPb max masked [H, W], oriented sigmoid maps [8, W, H] column-major)."""
    g = np.ascontiguousarray(gpb_maps, np.float64)
    _, H, W = g.shape
    gmax = np.empty((H, W), np.float32)
    gall = np.empty((8, W, H), np.float32)
    lib.orc_gpb_outputs(g.reshape(-1), np.ascontiguousarray(mpb_nmax, np.float64).reshape(-1), W, H,
                        gmax.reshape(-1), gall.reshape(-1))
    return gmax, gall


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def texton_features(textons: np.ndarray, superpixels: np.ndarray, nlabels: int) -> np.ndarray:
    """Per-superpixel texton histograms (SPFeatures::genTextonFeatures) -> [nlabels, 32]."""
    t = np.ascontiguousarray(textons, np.uint8)
    sp = np.ascontiguousarray(superpixels, np.int32)
    out = np.zeros((nlabels, 32), np.float64)
    if lib.orc_texton_features(_ptr(t), _ptr(sp), C.c_size_t(t.size), C.c_int(nlabels), _ptr(out)) != 0:
        raise ValueError("superpixel label outside [0, nlabels)")
    return out


def texton_patch_features(textons: np.ndarray, centres: np.ndarray) -> np.ndarray:
    """Texton histograms of the clipped 65 x 65 windows around centres [n, 2] (row, col) -> [n, 32]."""
    t = np.ascontiguousarray(textons, np.uint8)
    H, W = t.shape
    c = np.ascontiguousarray(centres, np.int32).reshape(-1, 2)
    out = np.zeros((c.shape[0], 32), np.float64)
    lib.orc_texton_patch_features(_ptr(t), C.c_int(W), C.c_int(H), _ptr(c), C.c_int(c.shape[0]), _ptr(out))
    return out


def pipeline(packed: np.ndarray, codebook: np.ndarray, K: int = 9, tol: float = 1e-10, m: int = 40):
    """Whole CPU path; returns a dict of every intermediate."""
    t = texton(packed, codebook)
    cues, mpb_all = local_cues(packed, t)
    nm = nonmax(mpb_all)
    Wd, D = build_w(nm)
    ev, vec, it = eigs(Wd, D, K=K, m=m, tol=tol)
    s = spb(vec, ev)
    g = gpb(cues, s)
    gmax, gall = gpb_outputs(g, nm)
    return dict(textons=t, cues=cues, mpb_all=mpb_all, mpb_nmax=nm, Wd=Wd, D=D, evals=ev, evecs=vec, iters=it,
                spb=s, gpb=g, gpb_max=gmax, gpb_all=gall)


# ---- the reference's own buildW kernels (oracle/_ref, built by `make ref`) ----
REF_BUILDW = Path(__file__).resolve().parent / "_ref" / "libref_buildw.so"


def ref_build_w(mpb_nmax: np.ndarray):
    """W (dense N x N) and its row sums from the reference's computeSupport +
    computeAffinities2 (buildW/src/ic.cc, affinity.cc, smatrix.cc) on the
    lattice spectralPb builds from mPb; None if _ref is not built."""
    if not REF_BUILDW.exists():
        return None
    import ctypes as C
    lib = C.CDLL(str(REF_BUILDW))
    f = lib.ref_build_w
    f.restype = C.c_int
    m = np.ascontiguousarray(mpb_nmax, np.float64)
    H, W = m.shape
    N = H * W
    cap = N * 81
    rows = np.empty(cap, np.int32)
    cols = np.empty(cap, np.int32)
    vals = np.empty(cap, np.float64)
    rsum = np.empty(N, np.float64)
    nnz = C.c_int(cap)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = f(C.c_int(W), C.c_int(H), P(m), C.byref(nnz), P(rows), P(cols), P(vals), P(rsum))
    if rc != 0:
        raise RuntimeError(f"ref_build_w failed ({rc})")
    n = nnz.value
    A = np.zeros((N, N))
    A[rows[:n], cols[:n]] = vals[:n]
    return A, rsum, (rows[:n].copy(), cols[:n].copy())


# ---- the reference's own lib_image (oracle/_ref, built by `make ref`) ----
REF_LIBIMAGE = Path(__file__).resolve().parent / "_ref" / "libref_libimage.so"
_REFL = None


def _refl():
    global _REFL
    if _REFL is None and REF_LIBIMAGE.exists():
        _REFL = C.CDLL(str(REF_LIBIMAGE))
    return _REFL


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def ref_gaussian(sigma: float) -> np.ndarray:
    """lib_image::gaussian(sigma) (lib_image.hh:404)."""
    lib = _refl()
    buf = np.empty(4096, np.float64)
    n = lib.ref_gaussian(C.c_double(sigma), _p(buf), C.c_int(buf.size))
    return buf[:n].copy()


def ref_lab_quantize(rgb: np.ndarray, nbins: int = 25) -> np.ndarray:
    """Gamma 2.5 / Lab / normalise / quantise through lib_image; rgb [..., 3] u8 -> bins [3, ...]."""
    lib = _refl()
    a = np.ascontiguousarray(rgb, np.uint8)
    n = a.size // 3
    out = np.empty((3, n), np.uint8)
    lib.ref_lab_quantize(_p(a), C.c_int(n), C.c_int(nbins), _p(out))
    return out.reshape((3,) + a.shape[:-1])


def ref_mirror(img: np.ndarray, border: int) -> np.ndarray:
    lib = _refl()
    m = np.ascontiguousarray(img, np.float64)
    H, W = m.shape
    out = np.empty((H + 2 * border, W + 2 * border), np.float64)
    if lib.ref_mirror(_p(m), C.c_int(W), C.c_int(H), C.c_int(border), _p(out)) != 0:
        raise ValueError("lib_image::border_mirror_2D rejected the arguments")
    return out


def ref_hist_gradient(labels: np.ndarray, r: int, ksigma: float) -> np.ndarray:
    """lib_image::hist_gradient_2D(labels, r, 8, gaussian(ksigma) or none) -> [8, H, W] (untrimmed)."""
    lib = _refl()
    lm = np.ascontiguousarray(labels, np.uint8)
    H, W = lm.shape
    out = np.empty((8, H, W), np.float64)
    lib.ref_hist_gradient(_p(lm), C.c_int(W), C.c_int(H), C.c_int(r), C.c_double(ksigma), _p(out))
    return out


def ref_nonmax(mpb_all: np.ndarray) -> np.ndarray:
    """nonmax_channels + lib_image::nonmax_oriented_2D + multiscalePb's clip."""
    lib = _refl()
    m = np.ascontiguousarray(mpb_all, np.float64)
    _, H, W = m.shape
    out = np.empty((H, W), np.float64)
    lib.ref_nonmax(_p(m), C.c_int(W), C.c_int(H), _p(out))
    return out


REF_FILTERS = Path(__file__).resolve().parent / "_ref" / "libref_filters.so"


def ref_spb_filter(ori: float) -> np.ndarray:
    """damascene gaussian_2D(1, 1, ori, deriv 1, support 3, 3) as 7 x 7 (x-major), FP32."""
    lib = C.CDLL(str(REF_FILTERS))
    f = np.empty(49, np.float32)
    lib.ref_spb_filter(C.c_double(ori), _p(f))
    return f.reshape(7, 7)

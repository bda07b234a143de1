"""ctypes binding of libgpb.so (the C ABI declared in include/gpb_capi.h).

Host-side numpy wrappers used by the tests and the benchmark.  The library is
the product; this module only marshals arrays.  Importing it fails loudly when
libgpb.so has not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libgpb.so"

GPB_OK, GPB_ERR_ARG, GPB_ERR_HIP, GPB_ERR_NOCONV, GPB_ERR_INTERNAL = 0, -1, -2, -3, -4
GPB_ERR_TOO_LARGE = -5

if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is missing: build it with `python ucm-mpi-gpu_amd/build.py`")
lib = C.CDLL(str(LIB_PATH))

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ctxp = C.c_void_p


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("gpb_ctx_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)])
_sig("gpb_ctx_destroy", None, [_ctxp])
_sig("gpb_ctx_last_error", C.c_char_p, [_ctxp])
_sig("gpb_device_count", C.c_int, [])
_sig("gpb_pack_rgb", None, [_u8p, C.c_size_t, _u32p])
_sig("gpb_synth_voronoi", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int, _u8p, C.c_void_p])
_sig("gpb_make_codebook", C.c_int, [_u32p, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, _f32p])
_sig("gpb_texton", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _f32p, _u8p])
_sig("gpb_texton_features", C.c_int,
     [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p])
_sig("gpb_texton_patch_features", C.c_int,
     [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p])
_sig("gpb_lab_quantize", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _u8p, _i32p])
_sig("gpb_local_cues", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _u8p, C.c_void_p, _f32p])
_sig("gpb_nonmax", C.c_int, [_ctxp, C.c_int, C.c_int, _f32p, _f32p])
_sig("gpb_set_eig_options", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int])
_sig("gpb_set_eig_convergence", C.c_int, [_ctxp, C.c_int])
_sig("gpb_set_lanczos_fusion", C.c_int, [_ctxp, C.c_int])
_sig("gpb_patch_eigenvectors", C.c_int,
     [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p])
_sig("gpb_build_w", C.c_int, [_ctxp, C.c_int, C.c_int, _f32p, _f32p, _f32p])
_sig("gpb_eigs", C.c_int, [_ctxp, C.c_int, C.c_int, _f32p, _f64p, C.c_void_p])
_sig("gpb_spectral_pb", C.c_int, [_ctxp, C.c_int, C.c_int, _f32p, _f32p, C.c_void_p])
_sig("gpb_spb_from_evecs", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, _f32p, _f64p, _f32p])
_sig("gpb_last_eigenvectors", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p])
_sig("gpb_compute", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _f32p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p])
_sig("gpb_compute_device", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p])
_sig("gpb_ic_program", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS"),
                                 np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS"), C.c_int])


class Stats(C.Structure):
    _fields_ = [("stage_ms", C.c_double * 6), ("lanczos_matvecs", C.c_int), ("lanczos_restarts", C.c_int),
                ("lanczos_basis", C.c_int), ("lanczos_block", C.c_int), ("converged", C.c_int),
                ("max_residual", C.c_double), ("evals", C.c_double * 32), ("fallback", C.c_int),
                ("reorth_second_passes", C.c_int), ("warm_mode", C.c_int), ("warm_ms", C.c_double),
                ("reorth_redone_steps", C.c_int)]


_sig("gpb_get_stats", C.c_int, [_ctxp, C.POINTER(Stats)])
_sig("gpb_set_profiling", C.c_int, [_ctxp, C.c_int])
_sig("gpb_set_profiling_classes", C.c_int, [_ctxp, C.c_uint])
_sig("gpb_set_block_size", C.c_int, [_ctxp, C.c_int])
_sig("gpb_set_spmv_tile", C.c_int, [_ctxp, C.c_int])
_sig("gpb_set_spmv_precision", C.c_int, [_ctxp, C.c_int])
_sig("gpb_set_reorth", C.c_int, [_ctxp, C.c_int])
_sig("gpb_set_lanczos_graph", C.c_int, [_ctxp, C.c_int])
_sig("gpb_set_block_fusion", C.c_int, [_ctxp, C.c_int])
_sig("gpb_set_reorth_threshold", C.c_int, [_ctxp, C.c_double])
_sig("gpb_set_verbose", C.c_int, [_ctxp, C.c_int])
_sig("gpb_compute_f64", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _f32p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p])
_sig("gpb_eigs_f64", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p])
_sig("gpb_cut_patches", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, _i32p, _i32p])
_sig("gpb_compute_patches", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _f32p, C.c_int, _i32p, _i32p, C.c_int,
                                      C.c_void_p, C.c_void_p])
_sig("gpb_compute_patches_device", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, _i32p, _i32p,
                                             C.c_int, C.c_void_p, C.c_void_p])
_sig("gpb_get_patch_stats", C.c_int, [_ctxp, C.c_int, C.c_void_p])
_sig("gpb_get_spmv_tile", C.c_int, [_ctxp, C.c_int, C.c_int, C.POINTER(C.c_int)])
_sig("gpb_apply_laplacian", C.c_int, [_ctxp, C.c_int, C.c_int, _f32p, C.c_int, C.c_int, _f32p, _f32p])
_sig("gpb_kernel_time", C.c_int, [_ctxp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long), C.POINTER(C.c_double)])
_sig("gpb_set_warm_start", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int])
_sig("gpb_warm_param_count", C.c_size_t, [])
_sig("gpb_warm_layer_info", C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)])
_sig("gpb_warm_init", C.c_int, [_ctxp, C.c_uint64])
_sig("gpb_warm_set_params", C.c_int, [_ctxp, _f32p])
_sig("gpb_warm_get_params", C.c_int, [_ctxp, _f32p])
_sig("gpb_warm_dataset_reset", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int])
_sig("gpb_warm_dataset_add", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _f32p, C.c_void_p])
_sig("gpb_warm_dataset_size", C.c_int, [_ctxp, C.POINTER(C.c_int)])
_sig("gpb_warm_features", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _f32p, _f32p])
_sig("gpb_warm_train", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64, C.c_int,
                                 C.c_void_p])
_sig("gpb_warm_eval_loss", C.c_int, [_ctxp, C.c_int, C.c_int, C.POINTER(C.c_double)])
_sig("gpb_warm_loss_grad", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p, C.POINTER(C.c_double),
                                     C.c_void_p, C.c_void_p])


def warm_param_count() -> int:
    return int(lib.gpb_warm_param_count())


def warm_layers() -> list:
    """Per layer: (cin, cout, relu, weight offset, bias offset) in the parameter vector."""
    out = []
    for l in range(12):
        ci, co, r = C.c_int(), C.c_int(), C.c_int()
        wo, bo = C.c_size_t(), C.c_size_t()
        _check(lib.gpb_warm_layer_info(l, C.byref(ci), C.byref(co), C.byref(r), C.byref(wo), C.byref(bo)),
               what="warm_layer_info")
        out.append((ci.value, co.value, r.value, wo.value, bo.value))
    return out
# computeGPb's hostGPbAllConcat (gpb_all): block i holds orientation
# GPB_ALL_BLOCK_ORIENTATION[i] (split angle o * pi / 8 of the half-disc cues),
# damascene's order (oriMap {4,5,6,7,0,1,2,3} over filter slots that hold
# orientation (3 - s) mod 8; damascene/src/damascene.cu:598-603,
# spectralPb/spectralPb.cu:101,706-712).
GPB_ALL_BLOCK_ORIENTATION = (7, 6, 5, 4, 3, 2, 1, 0)
KERNEL_CLASSES = ["texton", "cues", "nonmax", "affinity", "spmv", "dots", "update", "ritz", "spb", "gpb", "warm", "ucm"]


_sig("gpb_frame_size", C.c_size_t, [C.c_int, C.c_void_p])
_sig("gpb_frame_mats", C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t])
_sig("gpb_unframe_mats", C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p])

_CV_DEPTH = {np.dtype(np.uint8): 0, np.dtype(np.int8): 1, np.dtype(np.uint16): 2, np.dtype(np.int16): 3,
             np.dtype(np.int32): 4, np.dtype(np.float32): 5, np.dtype(np.float64): 6}


def frame_mats(mats) -> bytes:
    """The byte frame (count, then per matrix {rows, cols, channels, OpenCV type} and the raw elements) of
    2-D (rows, cols) or 3-D (rows, cols, channels) arrays: what MpiSendMatVec sends, in one buffer."""
    arrs = [np.ascontiguousarray(m) for m in mats]
    dims = np.zeros((max(1, len(arrs)), 4), np.int32)
    for i, a in enumerate(arrs):
        ch = a.shape[2] if a.ndim == 3 else 1
        dims[i] = (a.shape[0], a.shape[1], ch, _CV_DEPTH[a.dtype] + 8 * (ch - 1))
    size = lib.gpb_frame_size(len(arrs), dims.ctypes.data)
    if size == 0:
        raise GpbError(GPB_ERR_ARG, "frame_mats: unsupported matrix")
    ptrs = (C.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
    out = np.empty(size, np.uint8)
    _check(lib.gpb_frame_mats(len(arrs), dims.ctypes.data, ptrs, out.ctypes.data, size), what="frame_mats")
    return out.tobytes()


def unframe_mats(frame: bytes, max_count: int = 1024) -> list:
    """Inverse of frame_mats: the matrices of a frame, as arrays."""
    buf = np.frombuffer(frame, np.uint8)
    dims = np.zeros((max_count, 4), np.int32)
    offs = np.zeros(max_count, np.uint64)
    n = C.c_int(0)
    _check(lib.gpb_unframe_mats(buf.ctypes.data, buf.size, max_count, C.byref(n), dims.ctypes.data, offs.ctypes.data),
           what="unframe_mats")
    inv = {v: k for k, v in _CV_DEPTH.items()}
    out = []
    for i in range(n.value):
        r, c, ch, t = (int(v) for v in dims[i])
        dt = inv[t & 7]
        a = np.frombuffer(frame, dt, count=r * c * ch, offset=int(offs[i])).reshape((r, c, ch) if ch > 1 else (r, c))
        out.append(a.copy())
    return out


def ic_program(left: int, right: int, top: int, bottom: int):
    conn = np.zeros((81, 5), np.int8)
    sites = np.zeros((4096, 2), np.int8)
    rc = lib.gpb_ic_program(left, right, top, bottom, conn.reshape(-1), sites.reshape(-1), 4096)
    _check(rc, what="ic_program")
    start = (conn[:, 3].astype(np.int64) & 0x7f) | (conn[:, 4].astype(np.int64) << 7)
    return conn[:, 0].astype(bool), conn[:, 1].astype(bool), conn[:, 2].astype(np.int64), start, sites


class GpbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gpb status {code}: {msg}")
        self.code = code


def _check(rc: int, ctx=None, what: str = ""):
    if rc != GPB_OK:
        msg = lib.gpb_ctx_last_error(ctx).decode() if ctx else what
        raise GpbError(rc, msg)


def device_count() -> int:
    return lib.gpb_device_count()


# ---- OWT-UCM (host code, no context) ----
_sig("gpb_contours2ucm", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p])
_sig("gpb_ucm_stages", C.c_int, [C.c_int, C.c_int, C.c_void_p] + [C.c_void_p] * 6 + [C.c_void_p] * 3)
_sig("gpb_watershed8", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p])
_sig("gpb_contours2ucm_device", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p])
_sig("gpb_compute_ucm", C.c_int, [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p])
_sig("gpb_fit_contours", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p])


def contours2ucm(pb_oriented: np.ndarray, double_size: bool = True) -> np.ndarray:
    """Contours2Ucm of the oriented maps [8, H, W]: (2H+1, 2W+1) when double_size, else (H+1, W+1)."""
    pbo = np.ascontiguousarray(pb_oriented, np.float64)
    _, H, W = pbo.shape
    out = np.empty((2 * H + 1, 2 * W + 1) if double_size else (H + 1, W + 1))
    _check(lib.gpb_contours2ucm(W, H, pbo.ctypes.data, int(double_size), out.ctypes.data), what="contours2ucm")
    return out


def ucm_stages(pb_oriented: np.ndarray) -> dict:
    """The intermediate results of contours2ucm (see gpb_ucm_stages)."""
    pbo = np.ascontiguousarray(pb_oriented, np.float64)
    _, H, W = pbo.shape
    d = dict(pb8=np.empty((H, W), np.uint8), ws=np.empty((H, W), np.int32), ws_wt=np.empty((H, W)),
             ws_wt2=np.empty((2 * H + 1, 2 * W + 1)), labels=np.empty((H, W), np.int32),
             super_ucm=np.empty((2 * H + 1, 2 * W + 1)))
    reg, fail = C.c_int(), C.c_int()
    ms = np.zeros(7)
    _check(lib.gpb_ucm_stages(W, H, pbo.ctypes.data, d["pb8"].ctypes.data, d["ws"].ctypes.data, d["ws_wt"].ctypes.data,
                              d["ws_wt2"].ctypes.data, d["labels"].ctypes.data, d["super_ucm"].ctypes.data,
                              C.byref(reg), C.byref(fail), ms.ctypes.data), what="ucm_stages")
    d["regions"], d["contours_failed"] = reg.value, bool(fail.value)
    d["stage_ms"] = dict(zip(("pb8", "watershed", "contours", "strength", "grid_clean_regions", "merging",
                              "calibration"), ms.round(2).tolist()))
    return d


def watershed8(image: np.ndarray) -> np.ndarray:
    """8-connected watershed labels of an 8-bit image (0 = watershed pixel)."""
    img = np.ascontiguousarray(image, np.uint8)
    H, W = img.shape
    out = np.empty((H, W), np.int32)
    _check(lib.gpb_watershed8(W, H, img.ctypes.data, out.ctypes.data), what="watershed8")
    return out


def fit_contours(bw: np.ndarray):
    """Contours of a binary image: (vertices [nv, 2], edges [ne, 2], offsets [ne + 1], points [np, 2]), rows and
    columns; None where the subdivision cannot be completed."""
    b = np.ascontiguousarray(bw != 0, np.uint8)
    H, W = b.shape
    nv, ne, npt = C.c_int(), C.c_int(), C.c_int64()
    rc = lib.gpb_fit_contours(W, H, b.ctypes.data, 0, None, 0, None, None, 0, None, C.byref(nv), C.byref(ne),
                              C.byref(npt))
    if rc == GPB_ERR_INTERNAL:
        return None
    if rc not in (GPB_OK, GPB_ERR_TOO_LARGE):
        _check(rc, what="fit_contours")
    V = np.zeros((max(nv.value, 1), 2), np.int32)
    E = np.zeros((max(ne.value, 1), 2), np.int32)
    O = np.zeros(ne.value + 1, np.int64)
    P = np.zeros((max(npt.value, 1), 2), np.int32)
    _check(lib.gpb_fit_contours(W, H, b.ctypes.data, nv.value, V.ctypes.data, ne.value, E.ctypes.data, O.ctypes.data,
                                npt.value, P.ctypes.data, C.byref(nv), C.byref(ne), C.byref(npt)), what="fit_contours")
    return V[:nv.value], E[:ne.value], O, P[:npt.value]


def synth_voronoi(width: int, height: int, regions: int, noise_sigma: float, seed: int, classes: int = 1,
                  with_regions: bool = False):
    rgb = np.empty((height, width, 3), np.uint8)
    reg = np.empty((height, width), np.int32) if with_regions else None
    rc = lib.gpb_synth_voronoi(width, height, regions, noise_sigma, seed, classes, rgb,
                               reg.ctypes.data if reg is not None else None)
    _check(rc, what="synth_voronoi")
    return (rgb, reg) if with_regions else rgb


def pack_rgb(rgb: np.ndarray) -> np.ndarray:
    rgb = np.ascontiguousarray(rgb, np.uint8)
    out = np.empty(rgb.shape[:2], np.uint32)
    lib.gpb_pack_rgb(rgb.reshape(-1), rgb.shape[0] * rgb.shape[1], out.reshape(-1))
    return out


def cut_patches(width: int, height: int, grid: int, overlap: int, seed: int):
    """(rects, cores): [grid*grid, 4] int32 arrays of (x, y, w, h)."""
    rects = np.empty((grid * grid, 4), np.int32)
    cores = np.empty((grid * grid, 4), np.int32)
    rc = lib.gpb_cut_patches(width, height, grid, overlap, seed, rects.reshape(-1), cores.reshape(-1))
    if rc != GPB_OK:
        raise GpbError(rc, "bad patch grid")
    return rects, cores


def make_codebook(packed: np.ndarray, seed: int = 1, iters: int = 10, max_samples: int = 65536) -> np.ndarray:
    packed = np.ascontiguousarray(packed, np.uint32)
    cb = np.empty((13, 32), np.float32)
    rc = lib.gpb_make_codebook(packed.reshape(-1), packed.shape[1], packed.shape[0], seed, iters, max_samples,
                               cb.reshape(-1))
    _check(rc, what="make_codebook")
    return cb


class Context:
    """One HIP device + stream + workspace (gpb_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        self.nvec = 9
        rc = lib.gpb_ctx_create(device, C.byref(h))
        if rc != GPB_OK:
            raise GpbError(rc, f"cannot create a context on device {device}")
        self.h = h

    def close(self):
        if self.h:
            lib.gpb_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        _check(rc, self.h)

    def texton(self, packed: np.ndarray, codebook: np.ndarray) -> np.ndarray:
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        out = np.empty((H, W), np.uint8)
        self._chk(lib.gpb_texton(self.h, W, H, packed.reshape(-1),
                                 np.ascontiguousarray(codebook, np.float32).reshape(-1), out.reshape(-1)))
        return out

    def texton_features(self, packed: np.ndarray, codebook: np.ndarray, superpixels: np.ndarray,
                        nlabels: int) -> np.ndarray:
        """Per-superpixel texton histograms [nlabels, 32] (10 * count / size)."""
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        sp = np.ascontiguousarray(superpixels, np.int32)
        out = np.empty((nlabels, 32), np.float64)
        self._chk(lib.gpb_texton_features(self.h, W, H, packed.ctypes.data, np.ascontiguousarray(codebook, np.float32)
                                          .ctypes.data, sp.ctypes.data, nlabels, out.ctypes.data))
        return out

    def texton_patch_features(self, packed: np.ndarray, codebook: np.ndarray, centres: np.ndarray) -> np.ndarray:
        """Texton histograms [n, 32] of the clipped 65 x 65 windows around centres [n, 2] (row, col)."""
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        c = np.ascontiguousarray(centres, np.int32).reshape(-1, 2)
        out = np.empty((c.shape[0], 32), np.float64)
        self._chk(lib.gpb_texton_patch_features(self.h, W, H, packed.ctypes.data,
                                                np.ascontiguousarray(codebook, np.float32).ctypes.data, c.shape[0],
                                                c.ctypes.data, out.ctypes.data))
        return out

    def lab_quantize(self, packed: np.ndarray):
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        bins = np.empty((3, H, W), np.uint8)
        nb = np.empty(3, np.int32)
        self._chk(lib.gpb_lab_quantize(self.h, W, H, packed.reshape(-1), bins.reshape(-1), nb))
        return bins, nb

    def local_cues(self, packed: np.ndarray, textons: np.ndarray, with_cues: bool = True):
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        cues = np.empty((12, 8, H, W), np.float32) if with_cues else None
        mpb = np.empty((8, H, W), np.float32)
        self._chk(lib.gpb_local_cues(self.h, W, H, packed.reshape(-1), np.ascontiguousarray(textons, np.uint8).reshape(-1),
                                     cues.ctypes.data if cues is not None else None, mpb.reshape(-1)))
        return cues, mpb

    def nonmax(self, mpb_all: np.ndarray) -> np.ndarray:
        m = np.ascontiguousarray(mpb_all, np.float32)
        _, H, W = m.shape
        out = np.empty((H, W), np.float32)
        self._chk(lib.gpb_nonmax(self.h, W, H, m.reshape(-1), out.reshape(-1)))
        return out

    def set_eig_options(self, nvec=9, basis=48, keep=16, tol=1e-3, max_matvecs=20000):
        self._chk(lib.gpb_set_eig_options(self.h, nvec, basis, keep, tol, max_matvecs))
        self.nvec = nvec

    def set_lanczos_fusion(self, on: bool):
        """Fused single-vector Lanczos steps (default) or the separate kernels."""
        self._chk(lib.gpb_set_lanczos_fusion(self.h, 1 if on else 0))

    def set_eig_convergence(self, rel: bool):
        """False: residual <= tol; True: residual <= tol * Ritz value (Spectra's rule)."""
        self._chk(lib.gpb_set_eig_convergence(self.h, 1 if rel else 0))

    @staticmethod
    def _stats_dict(s) -> dict:
        return dict(stage_ms=list(s.stage_ms), matvecs=s.lanczos_matvecs, restarts=s.lanczos_restarts,
                    basis=s.lanczos_basis, block=s.lanczos_block, converged=s.converged, max_residual=s.max_residual,
                    evals=list(s.evals), fallback=s.fallback, second_passes=s.reorth_second_passes,
                    redone_steps=s.reorth_redone_steps)

    def patch_stats(self, patch: int) -> dict:
        s = Stats()
        self._chk(lib.gpb_get_patch_stats(self.h, patch, C.byref(s)))
        return self._stats_dict(s)

    def patch_eigenvectors(self, patch: int, w: int, h: int, nvec: int = 9):
        """Eigenvectors [nvec, h, w], degree [h, w], W planes [40, h, w] and mPb_nmax [h, w] of a patch
        of the last patch call (w, h: that patch's rectangle)."""
        ev = np.empty((nvec, h, w), np.float32)
        deg = np.empty((h, w), np.float32)
        wup = np.empty((40, h, w), np.float32)
        nm = np.empty((h, w), np.float32)
        self._chk(lib.gpb_patch_eigenvectors(self.h, patch, nvec, ev.ctypes.data, deg.ctypes.data, wup.ctypes.data,
                                             nm.ctypes.data))
        return ev, deg, wup, nm

    def compute_patches(self, packed: np.ndarray, codebook: np.ndarray, rects: np.ndarray, cores: np.ndarray,
                        mode: int = 1):
        """Patch-wise gPb (host buffers): returns stitched gpb_max [H,W] and gpb_orient [8,H,W]."""
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        r = np.ascontiguousarray(rects, np.int32)
        c = np.ascontiguousarray(cores, np.int32)
        gmax = np.zeros((H, W), np.float32)
        orient = np.zeros((8, H, W), np.float32)
        self._chk(lib.gpb_compute_patches(self.h, W, H, packed.reshape(-1), np.ascontiguousarray(codebook, np.float32),
                                          r.shape[0], r.reshape(-1), c.reshape(-1), mode, gmax.ctypes.data,
                                          orient.ctypes.data))
        return gmax, orient

    def compute_patches_device(self, W: int, H: int, d_packed: int, d_codebook: int, rects: np.ndarray,
                               cores: np.ndarray, mode: int, d_gmax: int = 0, d_orient: int = 0):
        r = np.ascontiguousarray(rects, np.int32)
        c = np.ascontiguousarray(cores, np.int32)
        self._chk(lib.gpb_compute_patches_device(self.h, W, H, d_packed, d_codebook, r.shape[0], r.reshape(-1),
                                                 c.reshape(-1), mode, d_gmax or None, d_orient or None))

    def stats(self) -> dict:
        s = Stats()
        self._chk(lib.gpb_get_stats(self.h, C.byref(s)))
        return dict(stage_ms=list(s.stage_ms), matvecs=s.lanczos_matvecs, restarts=s.lanczos_restarts,
                    basis=s.lanczos_basis, block=s.lanczos_block, converged=s.converged, max_residual=s.max_residual,
                    evals=list(s.evals), fallback=s.fallback,
                    second_passes=s.reorth_second_passes, warm_mode=s.warm_mode, warm_ms=s.warm_ms,
                    redone_steps=s.reorth_redone_steps)

    def build_w(self, mpb_nmax: np.ndarray):
        m = np.ascontiguousarray(mpb_nmax, np.float32)
        H, W = m.shape
        wu = np.empty((40, H, W), np.float32)
        D = np.empty((H, W), np.float32)
        self._chk(lib.gpb_build_w(self.h, W, H, m.reshape(-1), wu.reshape(-1), D.reshape(-1)))
        return wu, D

    def eigs(self, mpb_nmax: np.ndarray, with_vectors: bool = True):
        """The ctx's nvec smallest generalised eigenpairs (set_eig_options)."""
        m = np.ascontiguousarray(mpb_nmax, np.float32)
        H, W = m.shape
        nvec = self.nvec
        ev = np.empty(nvec, np.float64)
        vec = np.empty((nvec, H, W), np.float32) if with_vectors else None
        self._chk(lib.gpb_eigs(self.h, W, H, m.reshape(-1), ev, vec.ctypes.data if vec is not None else None))
        return ev, vec

    def spectral_pb(self, mpb_nmax: np.ndarray):
        """sPb [8,H,W] and the ctx's nvec eigenvalues (set_eig_options)."""
        m = np.ascontiguousarray(mpb_nmax, np.float32)
        H, W = m.shape
        nvec = self.nvec
        spb = np.empty((8, H, W), np.float32)
        ev = np.empty(nvec, np.float64)
        self._chk(lib.gpb_spectral_pb(self.h, W, H, m.reshape(-1), spb.reshape(-1), ev.ctypes.data))
        return spb, ev

    def spb_from_evecs(self, evecs: np.ndarray, evals: np.ndarray) -> np.ndarray:
        e = np.ascontiguousarray(evecs, np.float32)
        K, H, W = e.shape
        spb = np.empty((8, H, W), np.float32)
        self._chk(lib.gpb_spb_from_evecs(self.h, W, H, K, e.reshape(-1), np.ascontiguousarray(evals, np.float64),
                                         spb.reshape(-1)))
        return spb

    def compute(self, packed: np.ndarray, codebook: np.ndarray, orient: bool = True):
        """computeGPb semantics; returns dict of gpb_max [H,W], gpb_all [8,W,H], gpb_orient [8,H,W] (unless
        orient is False), mpb_nmax."""
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        gmax = np.empty((H, W), np.float32)
        gall = np.empty((8, W, H), np.float32)
        gor = np.empty((8, H, W), np.float32) if orient else None
        nm = np.empty((H, W), np.float32)
        self._chk(lib.gpb_compute(self.h, W, H, packed.reshape(-1), np.ascontiguousarray(codebook, np.float32).reshape(-1),
                                  gmax.ctypes.data, gall.ctypes.data, gor.ctypes.data if orient else None,
                                  nm.ctypes.data))
        out = dict(gpb_max=gmax, gpb_all=gall, mpb_nmax=nm)
        if orient:
            out["gpb_orient"] = gor
        return out

    def set_profiling(self, on: bool):
        self._chk(lib.gpb_set_profiling(self.h, int(on)))

    def set_profiling_classes(self, names):
        """Per-kernel events only around the named kernel classes (empty: off)."""
        mask = 0
        for n in names:
            mask |= 1 << KERNEL_CLASSES.index(n)
        self._chk(lib.gpb_set_profiling_classes(self.h, C.c_uint(mask)))

    def kernel_times(self) -> dict:
        out = {}
        for i, n in enumerate(KERNEL_CLASSES):
            ms, cnt, by = C.c_double(), C.c_long(), C.c_double()
            self._chk(lib.gpb_kernel_time(self.h, i, C.byref(ms), C.byref(cnt), C.byref(by)))
            out[n] = dict(ms=ms.value, launches=cnt.value, bytes=by.value)
        return out

    def compute_device(self, width: int, height: int, d_packed: int, d_codebook: int, d_gmax: int = 0, d_gall: int = 0,
                       d_orient: int = 0):
        self._chk(lib.gpb_compute_device(self.h, width, height, d_packed, d_codebook, d_gmax or None, d_gall or None,
                                         d_orient or None))

    def contours2ucm_device(self, d_pb_oriented: int, width: int, height: int, double_size: bool = True) -> np.ndarray:
        """OWT-UCM of oriented FP32 maps in device memory (pointer to [8, H, W]); the ucm on the host."""
        out = np.empty((2 * height + 1, 2 * width + 1) if double_size else (height + 1, width + 1))
        self._chk(lib.gpb_contours2ucm_device(self.h, width, height, d_pb_oriented, int(double_size), out.ctypes.data))
        return out

    def compute_ucm(self, width: int, height: int, d_packed: int, d_codebook: int, double_size: bool = True):
        """gPb of a device image, then OWT-UCM of its oriented maps; the ucm on the host."""
        out = np.empty((2 * height + 1, 2 * width + 1) if double_size else (height + 1, width + 1))
        self._chk(lib.gpb_compute_ucm(self.h, width, height, d_packed, d_codebook, int(double_size), out.ctypes.data))
        return out

    def last_eigenvectors(self, width: int, height: int, nvec: int = 9, with_degree: bool = True):
        """Eigenvectors [nvec, H, W] and degree [H, W] of the last compute / compute_device call."""
        vec = np.empty((nvec, height, width), np.float32)
        deg = np.empty((height, width), np.float32) if with_degree else None
        self._chk(lib.gpb_last_eigenvectors(self.h, width, height, nvec, vec.ctypes.data,
                                            deg.ctypes.data if with_degree else None))
        return vec, deg

    def set_block_size(self, block: int):
        self._chk(lib.gpb_set_block_size(self.h, block))

    # ---- learned warm start ----
    def set_warm_start(self, mode: int, basis: int = 0, keep: int = 0):
        """0 plain Lanczos, 1 learned block, 2 block Lanczos from a random block."""
        self._chk(lib.gpb_set_warm_start(self.h, mode, basis, keep))

    def warm_init(self, seed: int):
        self._chk(lib.gpb_warm_init(self.h, seed))

    def warm_params(self) -> np.ndarray:
        p = np.empty(warm_param_count(), np.float32)
        self._chk(lib.gpb_warm_get_params(self.h, p))
        return p

    def set_warm_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != warm_param_count():
            raise ValueError("parameter vector has the wrong length")
        self._chk(lib.gpb_warm_set_params(self.h, p))

    def warm_dataset_reset(self, width: int, height: int, capacity: int):
        self._chk(lib.gpb_warm_dataset_reset(self.h, width, height, capacity))

    def warm_dataset_add(self, packed: np.ndarray, codebook: np.ndarray) -> np.ndarray:
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        ev = np.empty(self.nvec)
        self._chk(lib.gpb_warm_dataset_add(self.h, W, H, packed, np.ascontiguousarray(codebook, np.float32),
                                           ev.ctypes.data))
        return ev

    def warm_features(self, packed: np.ndarray, codebook: np.ndarray) -> np.ndarray:
        """Cue features [32, ceil(H/4), ceil(W/4)] (4 cues x 8 orientations, 4x4 cell means)."""
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        f = np.empty((32, (H + 3) // 4, (W + 3) // 4), np.float32)
        self._chk(lib.gpb_warm_features(self.h, W, H, packed, np.ascontiguousarray(codebook, np.float32),
                                        f.reshape(-1)))
        return f

    def warm_dataset_size(self) -> int:
        n = C.c_int()
        self._chk(lib.gpb_warm_dataset_size(self.h, C.byref(n)))
        return n.value

    def warm_train(self, first: int, count: int, epochs: int, batch: int, lr: float, seed: int,
                   augment: bool = False) -> np.ndarray:
        """Trains on training-set entries [first, first + count) (augment: random grid symmetries);
        returns the per-epoch mean loss."""
        loss = np.empty(epochs, np.float32)
        self._chk(lib.gpb_warm_train(self.h, first, count, epochs, batch, lr, seed, int(augment), loss.ctypes.data))
        return loss

    def warm_eval_loss(self, first: int, count: int) -> float:
        v = C.c_double()
        self._chk(lib.gpb_warm_eval_loss(self.h, first, count, C.byref(v)))
        return v.value

    def warm_loss_grad(self, x: np.ndarray, targets: np.ndarray, dweights: np.ndarray):
        """x [n,hq,wq,34], targets [n,hq,wq,8], dweights [n,hq,wq] -> (loss, grad, y [n,hq,wq,8])."""
        x = np.ascontiguousarray(x, np.float32)
        n, hq, wq, _ = x.shape
        g = np.empty(warm_param_count(), np.float32)
        y = np.empty((n, hq, wq, 8), np.float32)
        v = C.c_double()
        self._chk(lib.gpb_warm_loss_grad(self.h, n, hq, wq, x.reshape(-1),
                                         np.ascontiguousarray(targets, np.float32).reshape(-1),
                                         np.ascontiguousarray(dweights, np.float32).reshape(-1), C.byref(v),
                                         g.ctypes.data, y.ctypes.data))
        return v.value, g, y

    SPMV_TILES = (42, 44, 48, 82, 84, 161, 162)

    def set_spmv_tile(self, tile: int):
        """SpMV tile code 10 * wavefronts + rows per thread (one of SPMV_TILES), or 0 to autotune per shape."""
        self._chk(lib.gpb_set_spmv_tile(self.h, tile))

    def compute_f64(self, packed: np.ndarray, codebook: np.ndarray, with_evecs: bool = True) -> dict:
        """The all-FP64 GPU path: mpb_nmax, gpb_orient [8,H,W], gpb_max, evals, evecs [nvec,H,W] (float64)."""
        packed = np.ascontiguousarray(packed, np.uint32)
        H, W = packed.shape
        K = self.nvec
        out = dict(mpb_nmax=np.empty((H, W)), gpb_orient=np.empty((8, H, W)), gpb_max=np.empty((H, W)),
                   evals=np.empty(K), evecs=np.empty((K, H, W)) if with_evecs else None)
        self._chk(lib.gpb_compute_f64(self.h, W, H, packed.reshape(-1), np.ascontiguousarray(codebook, np.float32),
                                      out["mpb_nmax"].ctypes.data, out["gpb_orient"].ctypes.data,
                                      out["gpb_max"].ctypes.data, out["evals"].ctypes.data,
                                      out["evecs"].ctypes.data if with_evecs else None))
        return out

    def eigs_f64(self, mpb_nmax: np.ndarray, with_vectors: bool = True):
        m = np.ascontiguousarray(mpb_nmax, np.float64)
        H, W = m.shape
        ev = np.empty(self.nvec)
        vec = np.empty((self.nvec, H, W)) if with_vectors else None
        self._chk(lib.gpb_eigs_f64(self.h, W, H, m.ctypes.data, ev.ctypes.data,
                                   vec.ctypes.data if vec is not None else None))
        return ev, vec

    def set_verbose(self, level: int):
        self._chk(lib.gpb_set_verbose(self.h, level))

    def set_spmv_fp64(self, on: bool):
        """SpMV arithmetic in FP64 (default) or FP32; HBM operands are FP32 either way."""
        self._chk(lib.gpb_set_spmv_precision(self.h, 1 if on else 0))

    def set_block_fusion(self, mode: int):
        """Block Lanczos passes: 0 separate, 1 fused (same sums), 2 fused with the DGKS-gated second pass (default)."""
        self._chk(lib.gpb_set_block_fusion(self.h, int(mode)))

    def set_lanczos_graph(self, on: bool):
        """Replay single-vector Lanczos restart cycles from hipGraphs (default) or launch directly."""
        self._chk(lib.gpb_set_lanczos_graph(self.h, 1 if on else 0))

    def set_reorth(self, partial):
        """Partial reorthogonalisation (True: the default, optimistic cycles; 1: gated launches every step) or
        the whole-basis pass every step (False / 0)."""
        mode = (2 if partial else 0) if isinstance(partial, bool) else int(partial)
        self._chk(lib.gpb_set_reorth(self.h, mode))

    def set_reorth_threshold(self, level: float):
        """Trigger level of the omega recurrence and of the block solver's semi-orthogonality gate
        (0 = sqrt(2^-24), the default)."""
        self._chk(lib.gpb_set_reorth_threshold(self.h, C.c_double(level)))

    def spmv_tile(self, width: int, height: int) -> int:
        r = C.c_int(0)
        self._chk(lib.gpb_get_spmv_tile(self.h, width, height, C.byref(r)))
        return r.value

    def apply_laplacian(self, mpb_nmax: np.ndarray, x: np.ndarray, block: int = 1) -> np.ndarray:
        m = np.ascontiguousarray(mpb_nmax, np.float32)
        xx = np.ascontiguousarray(x, np.float32)
        k, H, W = xx.shape
        y = np.empty_like(xx)
        self._chk(lib.gpb_apply_laplacian(self.h, W, H, m.reshape(-1), k, block, xx.reshape(-1), y.reshape(-1)))
        return y

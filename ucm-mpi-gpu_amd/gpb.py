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
_sig("gpb_lab_quantize", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _u8p, _i32p])
_sig("gpb_local_cues", C.c_int, [_ctxp, C.c_int, C.c_int, _u32p, _u8p, C.c_void_p, _f32p])
_sig("gpb_nonmax", C.c_int, [_ctxp, C.c_int, C.c_int, _f32p, _f32p])


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

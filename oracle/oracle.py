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


def local_cues(packed: np.ndarray, textons: np.ndarray):
    """Returns cues[12, 8, H, W] and mpb_all[8, H, W] (FP64)."""
    packed = np.ascontiguousarray(packed, np.uint32)
    H, W = packed.shape
    cues = np.empty((12, 8, H, W), np.float64)
    mpb = np.empty((8, H, W), np.float64)
    lib.orc_local_cues(packed.reshape(-1), np.ascontiguousarray(textons, np.uint8).reshape(-1), W, H,
                       cues.reshape(-1), mpb.ctypes.data)
    return cues, mpb


def nonmax(mpb_all: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(mpb_all, np.float64)
    _, H, W = m.shape
    out = np.empty((H, W), np.float64)
    lib.orc_nonmax(m.reshape(-1), W, H, out.reshape(-1))
    return out

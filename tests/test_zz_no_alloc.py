"""gpb_nonmax_noalloc, the entry point that allocates nothing: through the
C ABI on device buffers the caller owns, its output equals gpb_nonmax's bit
for bit on two calls in a row, the second on a smaller image (the scratch
array keeps the first call's contents, which the second must not see); an
image beyond GPB_NOALLOC_MAX_PIXELS is rejected with GPB_ERR_TOO_LARGE
before any launch, and bad arguments with GPB_ERR_ARG."""
import ctypes

import numpy as np
import pytest

from conftest import synth_packed

GPB_ERR_ARG, GPB_ERR_TOO_LARGE = -1, -5
MAX_PIXELS = 2048 * 2048


def _entry(gpb):
    f = gpb.lib.gpb_nonmax_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return f


class _Hip:
    """The caller's side: device buffers from the HIP runtime that libgpb.so
    itself links (the caller allocates; the entry point does not)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipSetDevice"):
            getattr(self.lib, name).restype = ctypes.c_int
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert self.lib.hipSetDevice(0) == 0

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p.value

    def h2d(self, dst, arr):
        assert self.lib.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.lib.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0


@pytest.mark.gpu
def test_nonmax_noalloc_matches_nonmax_gpu(gpb, ctx):
    f = _entry(gpb)
    hip = _Hip()
    cap = 640 * 480
    d_in, d_out = hip.malloc(8 * cap * 4), hip.malloc(cap * 4)
    try:
        for W, H, seed in [(640, 480, 1), (200, 150, 2)]:
            print(f"nonmax_noalloc {W}x{H} seed {seed}", flush=True)
            rgb, packed = synth_packed(gpb, W, H, regions=12, seed=seed)
            cb = gpb.make_codebook(packed, seed=seed)
            _, mpb_all = ctx.local_cues(packed, ctx.texton(packed, cb), with_cues=False)
            # plus exact ties between orientations and negative values, which
            # the argmax and the clip must treat as gpb_nonmax does
            mpb_all[:, :8, :8] = 0.25
            mpb_all[3, 8:12, :] = -0.5
            want = ctx.nonmax(mpb_all)
            hip.h2d(d_in, np.ascontiguousarray(mpb_all, np.float32))
            hip.h2d(d_out, np.full(W * H, -7.0, np.float32))
            hip.sync()
            rc = f(W, H, d_in, d_out, None)  # the null stream
            assert rc == 0, rc
            hip.sync()
            got = np.empty((H, W), np.float32)
            hip.d2h(got, d_out)
            assert np.array_equal(got, want), np.abs(got - want).max()
        # rejected before any launch: too many pixels, a grid too tall, bad arguments
        assert f(2049, 2048, d_in, d_out, None) == GPB_ERR_TOO_LARGE
        assert f(MAX_PIXELS // 65536 + 1, 65536, d_in, d_out, None) in (GPB_ERR_ARG, GPB_ERR_TOO_LARGE)
        assert f(0, 10, d_in, d_out, None) == GPB_ERR_ARG
        assert f(10, 10, None, d_out, None) == GPB_ERR_ARG
        hip.sync()
    finally:
        hip.lib.hipFree(d_in)
        hip.lib.hipFree(d_out)

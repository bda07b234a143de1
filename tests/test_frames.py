"""Byte framing of the reference's matrix messages (MpiSendMat / MpiRecvMat,
MpiSendMatVec / MpiRecvMatVec, src/utilities.cpp:382-460), for moving gPb
outputs between processes without MPI.

Known answers: a frame is the int32 count, then per matrix the four int32
header words {rows, cols, channels, OpenCV type} the reference sends first
(type = depth + 8 (channels - 1); CV_64FC1 = 6, CV_32FC1 = 5, CV_8UC3 = 16) and
the raw row-major elements it sends second, with no padding.  Round trips are
exact for every depth; truncated, overlong and inconsistent frames are
rejected."""
import struct

import numpy as np
import pytest


def test_frame_layout_known_answer(gpb):
    a = np.arange(6, dtype=np.float64).reshape(2, 3) / 7
    b = np.arange(24, dtype=np.uint8).reshape(2, 4, 3)
    f = gpb.frame_mats([a, b])
    assert f[:4] == struct.pack("<i", 2)
    assert f[4:20] == struct.pack("<4i", 2, 3, 1, 6)
    assert f[20:68] == a.tobytes()
    assert f[68:84] == struct.pack("<4i", 2, 4, 3, 16)
    assert f[84:] == b.tobytes() and len(f) == 84 + 24


def test_frame_round_trip(gpb):
    rng = np.random.default_rng(3)
    mats = [rng.random((5, 7)), rng.random((3, 2)).astype(np.float32), rng.integers(0, 255, (4, 4, 3), np.uint8),
            rng.integers(-9, 9, (2, 9), np.int16), rng.integers(0, 60000, (3, 3), np.uint16),
            rng.integers(-5, 5, (1, 4), np.int32), np.zeros((0, 5))]
    back = gpb.unframe_mats(gpb.frame_mats(mats))
    assert len(back) == len(mats)
    for m, r in zip(mats, back):
        assert m.dtype == r.dtype and m.shape == r.shape and np.array_equal(m, r)
    assert gpb.unframe_mats(gpb.frame_mats([])) == []


def test_frame_rejects_malformed(gpb):
    f = gpb.frame_mats([np.ones((3, 3))])
    for bad in (f[:-1], f + b"\0", f[:10], struct.pack("<i", -1)):
        with pytest.raises(gpb.GpbError):
            gpb.unframe_mats(bad)
    # a header whose type disagrees with its channel count
    hdr = bytearray(f)
    hdr[12:16] = struct.pack("<i", 3)
    with pytest.raises(gpb.GpbError):
        gpb.unframe_mats(bytes(hdr))
    with pytest.raises(gpb.GpbError):
        gpb.unframe_mats(gpb.frame_mats([np.ones((2, 2))] * 3), max_count=2)

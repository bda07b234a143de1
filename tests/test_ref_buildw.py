"""The oracle's intervening-contour affinity against the reference's own
buildW kernels (buildW/src/ic.cc, affinity.cc, smatrix.cc compiled from the
reference tree into oracle/_ref by `make -C oracle ref`).

This pins the whole of stage (c) on the reference's code: the Bresenham
walks, the exact-approximant predicate (ic.cc:249-259), the corner
injections (ic.cc:427-457, 481-), C_IC_SS (affinity.cc:40-52), the disc, the
symmetrisation (smatrix.cc:95-117) and the row sums that buildW exports as
valD (buildW.cpp:160-183): the oracle's W and row sums must equal the
reference's bit for bit, with the same sparsity pattern.  Skipped where the reference tree is absent (no _ref).
"""
import numpy as np
import pytest

from conftest import synth_packed


def _cases(gpb, orc):
    rng = np.random.default_rng(5)
    out = []
    for H, W in [(6, 7), (12, 16), (19, 23)]:
        m = rng.random((H, W)) * (rng.random((H, W)) < 0.35)
        out.append((f"random {W}x{H}", m))
    ridge = np.zeros((17, 15))
    ridge[:, 7] = 1.0
    ridge[8, :] = 0.6
    out.append(("ridge cross 15x17", ridge))
    _, packed = synth_packed(gpb, 30, 26, regions=6, seed=9)
    cb = gpb.make_codebook(packed, seed=9)
    t = orc.texton(packed, cb)
    _, mpb = orc.local_cues(packed, t)
    out.append(("pipeline mPb 30x26", orc.nonmax(mpb)))
    out.append(("zero 9x8", np.zeros((8, 9))))
    return out


def test_oracle_w_equals_reference(gpb, orc):
    if not orc.REF_BUILDW.exists():
        pytest.skip("reference buildW not built (oracle/_ref needs the reference tree)")
    for name, m in _cases(gpb, orc):
        A_ref, rsum_ref, (rows, cols) = orc.ref_build_w(m)
        Wd, D = orc.build_w(m)
        A = orc.dense_w(Wd)
        assert np.array_equal(A_ref != 0, A != 0), name  # same sparsity (disc clipped at the border)
        np.testing.assert_array_equal(A, A_ref, err_msg=name)
        np.testing.assert_array_equal(D.reshape(-1), rsum_ref, err_msg=name)
        assert np.array_equal(A_ref, A_ref.T), name  # the reference's symmetrize
        # buildW's scanline numbering: every row lists itself with W_pp = 1
        assert np.all(np.diag(A_ref) == 1.0), name

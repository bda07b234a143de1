"""The reference's own CPU gPb path on one image patch, timed per stage: the
CPU baseline of bench.py.  Test infrastructure (like the rest of oracle/):
only tests/ and bench.py's cpu_baseline leg use it, never the product.

Stages and whose code runs them:
  (a) textons     the oracle's port (SPFeatures' Schmid path needs OpenCV,
                  which this image lacks: src/SPFeatures.cpp:277-360);
  (b) cues        the reference's lib_image (oracle/_ref/libref_libimage.so):
                  border_mirror_2D, Lab quantisation and hist_gradient_2D for
                  the 12 cues as src/mex_pb_parts_final_selected.cpp:100-227
                  calls them, the mPb sum of src/multiscalePb.cpp:216-245, and
                  nonmax_channels + nonmax_oriented_2D + the clip;
  (c) W           the reference's buildW kernels (oracle/_ref/libref_buildw.so:
                  computeSupport, computeAffinities2, symmetrize);
  (d) eigensolve  the reference's Spectra SymGEigsSolver
                  (oracle/_ref/libref_spectra.so) as src/spectralPb.cpp:196-244
                  sets it up, compute() with its defaults (tol 1e-10, 1000
                  restarts); nev = 9 eigenpairs as the configuration asks
                  (the reference's CPU path uses 17, src/globalPb.cpp:65),
                  ncv = 2 nev as the reference does;
  (e) sPb, gPb    the oracle's port (oeFilter reads uninitialised memory,
                  src/oeFilter.cpp:145-154).
Everything runs in this process on one thread.
"""
from __future__ import annotations

import ctypes as C
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = HERE / "_ref"
MPB_W = np.array([0.0146, 0.0145, 0.0163, 0.0210, 0.0243, 0.0287, 0.0166, 0.0185, 0.0204, 0.0101, 0.0111,
                  0.0141])  # src/multiscalePb.cpp:18-26


def available() -> bool:
    return all((REF / n).exists() for n in ("libref_libimage.so", "libref_buildw.so", "libref_spectra.so"))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def build_w_triplets(mpb_nmax: np.ndarray):
    """The reference's W (all rows, diagonal included) as triplets, and its row sums."""
    lib = C.CDLL(str(REF / "libref_buildw.so"))
    f = lib.ref_build_w
    f.restype = C.c_int
    m = np.ascontiguousarray(mpb_nmax, np.float64)
    H, W = m.shape
    N = H * W
    cap = N * 81
    rows, cols = np.empty(cap, np.int32), np.empty(cap, np.int32)
    vals, rsum = np.empty(cap, np.float64), np.empty(N, np.float64)
    nnz = C.c_int(cap)
    rc = f(C.c_int(W), C.c_int(H), _p(m), C.byref(nnz), _p(rows), _p(cols), _p(vals), _p(rsum))
    if rc != 0:
        raise RuntimeError(f"ref_build_w failed ({rc})")
    n = nnz.value
    return rows[:n], cols[:n], vals[:n], rsum


def spectra_geigs(rows, cols, vals, deg, nev=9, ncv=18, maxit=1000, tol=1e-10):
    """Spectra SymGEigsSolver on (D - W, D): (evals ascending, evecs [nev][n], converged, iterations,
    operations); pairs Spectra did not converge are NaN."""
    lib = C.CDLL(str(REF / "libref_spectra.so"))
    f = lib.ref_spectra_geigs
    f.restype = C.c_int
    n = deg.size
    ev = np.full(nev, np.nan)
    vec = np.full((nev, n), np.nan)
    it, ops = C.c_int(0), C.c_int(0)
    r = np.ascontiguousarray(rows, np.int32)
    c = np.ascontiguousarray(cols, np.int32)
    v = np.ascontiguousarray(vals, np.float64)
    d = np.ascontiguousarray(deg, np.float64)
    nconv = f(C.c_int(n), C.c_int(r.size), _p(r), _p(c), _p(v), _p(d), C.c_int(nev), C.c_int(ncv), C.c_int(maxit),
              C.c_double(tol), _p(ev), _p(vec), C.byref(it), C.byref(ops))
    return ev, vec, nconv, it.value, ops.value


def run_patch(orc, packed: np.ndarray, codebook: np.ndarray, nev: int = 9, ncv: int = 18, tol: float = 1e-10,
              maxit: int = 1000) -> dict:
    """The whole path on one packed patch; returns stage seconds, solver counts and the outputs."""
    H, W = packed.shape
    t = {}
    t0 = time.perf_counter()
    tx = orc.texton(packed, codebook)
    t["texton"] = time.perf_counter() - t0

    t0 = time.perf_counter()
    B = 30
    rgb = np.stack([(packed >> s) & 0xFF for s in (0, 8, 16)], axis=-1).astype(np.float64)
    rgbm = np.stack([orc.ref_mirror(rgb[..., ch], B) for ch in range(3)], axis=-1).astype(np.uint8)
    bins = orc.ref_lab_quantize(rgbm, 25)
    txm = orc.ref_mirror(tx.astype(np.float64), B).astype(np.uint8)
    cues = np.empty((12, 8, H, W))
    for s, (rb, rc) in enumerate(zip((3, 5, 10), (5, 10, 20))):
        cues[0 + s] = orc.ref_hist_gradient(bins[0], rb, 0.1 * 25)[:, B:-B, B:-B]
        cues[3 + s] = orc.ref_hist_gradient(bins[1], rc, 0.05 * 25)[:, B:-B, B:-B]
        cues[6 + s] = orc.ref_hist_gradient(bins[2], rc, 0.05 * 25)[:, B:-B, B:-B]
        cues[9 + s] = orc.ref_hist_gradient(txm, rc, 0.0)[:, B:-B, B:-B]
    mpb_all = np.tensordot(MPB_W, cues, axes=(0, 0))
    nm = orc.ref_nonmax(mpb_all)
    t["cues"] = time.perf_counter() - t0

    t0 = time.perf_counter()
    rows, cols, vals, deg = build_w_triplets(nm)
    t["affinity"] = time.perf_counter() - t0

    t0 = time.perf_counter()
    ev, vec, nconv, iters, ops = spectra_geigs(rows, cols, vals, deg, nev=nev, ncv=ncv, maxit=maxit, tol=tol)
    t["eigensolve"] = time.perf_counter() - t0

    t0 = time.perf_counter()
    s = orc.spb(vec.reshape(nev, H, W), ev)
    g = orc.gpb(cues, s)
    gmax, gall = orc.gpb_outputs(g, nm)
    t["spb_gpb"] = time.perf_counter() - t0
    return dict(seconds=t, total=sum(t.values()), converged=nconv, iterations=iters, operations=ops, evals=ev,
                mpb_nmax=nm, gpb_max=gmax, deg=deg, evecs=vec)


def _worker(args):
    packed, codebook = args
    import importlib.util
    import sys

    name = "gpb_oracle_binding"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, HERE / "oracle.py")
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    r = run_patch(sys.modules[name], packed, codebook)
    return dict(seconds=r["seconds"], total=r["total"], converged=r["converged"], operations=r["operations"])


def main():
    """python oracle/ref_cpu_path.py --input patches.npz --workers P --out result.json

    patches.npz holds `patches` [n][h][w] uint32 (packed RGB) and `codebook`
    [13][32].  Runs patch 0 alone (one thread), then the first P patches at
    once in P forked processes (one patch per process, as the reference runs
    one MPI rank per patch), and writes the timings as JSON."""
    import argparse
    import json
    import multiprocessing as mp
    import os

    ap = argparse.ArgumentParser()
    ap.add_argument("--input", required=True)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    d = np.load(a.input)
    patches, cb = d["patches"], d["codebook"]
    P = max(1, min(a.workers, len(patches)))
    t0 = time.perf_counter()
    single = _worker((patches[0], cb))
    t_single = time.perf_counter() - t0
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(P) as pool:
        many = pool.map(_worker, [(patches[i], cb) for i in range(P)], chunksize=1)
    wall = time.perf_counter() - t0
    h, w = patches.shape[1:]
    out = dict(patch=[int(h), int(w)], single_seconds=t_single, single_stages=single["seconds"],
               single_converged=single["converged"], single_operations=single["operations"], workers=P,
               parallel_wall_seconds=wall, parallel_patch_seconds=[m["total"] for m in many],
               parallel_converged=[m["converged"] for m in many], nproc=os.cpu_count(),
               affinity=len(os.sched_getaffinity(0)))
    with open(a.out, "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()

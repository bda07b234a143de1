"""Evaluate the learned warm start against plain Lanczos on unseen images.

For each image (default 20 seeded 4096 x 4096 three-class Voronoi images,
1600 regions, sigma 8, seeds 5000..5019, codebook from 65536 sampled
responses) the whole gPb path runs once per eigensolve start, with the packed
image, codebook and outputs resident in HBM (the bench's convention): plain
single-vector Lanczos (mode 0), block Lanczos from the network's block
(mode 1) and block Lanczos from a random block (mode 2, the same solver
without the network).  All at the same residual target (1e-3); each run
reports wall time, eigensolve-stage time, network time, operator
applications, fallback and measured residual, and for the modes other than
plain the largest principal angle between their 8 nontrivial eigenvectors and
the plain solve's (D-weighted, read back after the timed run).  The first
image's runs are repeated once untimed so that allocation and tuning stay out
of the timings.  Prints one JSON object.

    python ucm-mpi-gpu_amd/eval_warm_start.py --weights ucm-mpi-gpu_amd/weights/warm_start.f32
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent


def _binding():
    name = "gpb_amd_binding"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG / "gpb.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def subspace_angle(Va: np.ndarray, Vb: np.ndarray, D: np.ndarray) -> float:
    """Largest principal angle between span(Va) and span(Vb) in the D inner product."""
    s = np.sqrt(D.astype(np.float64)).reshape(-1)
    A = Va.reshape(len(Va), -1).astype(np.float64) * s
    B = Vb.reshape(len(Vb), -1).astype(np.float64) * s
    La = np.linalg.cholesky(A @ A.T)
    Lb = np.linalg.cholesky(B @ B.T)
    M = np.linalg.solve(La, A @ B.T)
    M = np.linalg.solve(Lb, M.T).T  # Qa Qb^T with Qa = La^-1 A, Qb = Lb^-1 B orthonormal rows
    sv = np.linalg.svd(M, compute_uv=False)
    return float(np.arccos(np.clip(sv.min(), -1.0, 1.0)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", required=True)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--regions", type=int, default=1600)
    ap.add_argument("--seed0", type=int, default=5000)
    ap.add_argument("--count", type=int, default=20)
    ap.add_argument("--modes", default="0,1,2")
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--basis", type=int, default=80)
    ap.add_argument("--keep", type=int, default=24)
    ap.add_argument("--plain-basis", type=int, default=48)
    ap.add_argument("--plain-keep", type=int, default=16)
    args = ap.parse_args()
    import torch  # device buffers; torch's runtime is initialised before the library's context

    S = args.size
    dev = torch.device("cuda:0")
    d_gmax = torch.empty((S, S), dtype=torch.float32, device=dev)
    d_gall = torch.empty((8, S, S), dtype=torch.float32, device=dev)
    gpb = _binding()
    modes = [int(m) for m in args.modes.split(",")]
    ctx = gpb.Context(0)
    ctx.set_eig_options(nvec=9, basis=args.plain_basis, keep=args.plain_keep, tol=args.tol, max_matvecs=40000)
    ctx.set_warm_params(np.fromfile(args.weights, dtype="<f4"))
    names = {0: "plain", 1: "learned", 2: "random_block"}
    per_image = []
    for i in range(args.count):
        seed = args.seed0 + i
        rgb = gpb.synth_voronoi(S, S, args.regions, 8.0, seed, classes=3)
        packed = gpb.pack_rgb(rgb)
        del rgb
        cb = gpb.make_codebook(packed, seed=seed, iters=10)
        d_rgb = torch.from_numpy(packed.view(np.int32)).to(dev)
        d_cb = torch.from_numpy(cb).to(dev)
        torch.cuda.synchronize()
        row = dict(seed=seed)
        vecs = {}
        for rep in range(2 if i == 0 else 1):
            for m in modes:
                ctx.set_warm_start(m, args.basis, args.keep)
                torch.cuda.synchronize()
                t = time.perf_counter()
                ctx.compute_device(S, S, d_rgb.data_ptr(), d_cb.data_ptr(), d_gmax.data_ptr(), d_gall.data_ptr())
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t) * 1e3
                st = ctx.stats()
                row[names[m]] = dict(ms=round(ms, 1), eig_ms=round(st["stage_ms"][4], 1),
                                     warm_ms=round(st["warm_ms"], 2), matvecs=st["matvecs"],
                                     fallback=st["fallback"], residual=st["max_residual"],
                                     evals=[float(x) for x in st["evals"][:9]])
                if rep == (1 if i == 0 else 0):
                    v, deg = ctx.last_eigenvectors(S, S, 9, with_degree=(m == modes[0]))
                    vecs[m] = v[1:]
                    if deg is not None:
                        vecs["D"] = deg
        if 0 in vecs:
            for m in modes:
                if m != 0:
                    row[names[m]]["angle_vs_plain"] = subspace_angle(vecs[m], vecs[0], vecs["D"])
        del vecs
        ctx.set_warm_start(0)
        per_image.append(row)
        brief = {names[m]: (row[names[m]]["ms"], row[names[m]]["matvecs"], row[names[m]]["fallback"],
                            round(row[names[m]].get("angle_vs_plain", 0.0), 4)) for m in modes}
        print(f"image {i + 1}/{args.count} seed {seed}: (ms, matvecs, fallback, angle vs plain) {brief}",
              file=sys.stderr, flush=True)
    summary = {}
    for m in modes:
        k = names[m]
        summary[k] = dict(mean_ms=float(np.mean([r[k]["ms"] for r in per_image])),
                          mean_eig_ms=float(np.mean([r[k]["eig_ms"] for r in per_image])),
                          mean_matvecs=float(np.mean([r[k]["matvecs"] for r in per_image])),
                          fallbacks=int(sum(r[k]["fallback"] for r in per_image)),
                          max_residual=float(max(r[k]["residual"] for r in per_image)),
                          mean_warm_ms=float(np.mean([r[k]["warm_ms"] for r in per_image])))
    for m in modes:
        if m == 0 or 0 not in modes:
            continue
        k = names[m]
        sp = [r["plain"]["ms"] / r[k]["ms"] for r in per_image]
        se = [r["plain"]["eig_ms"] / r[k]["eig_ms"] for r in per_image]
        summary[k]["speedup_vs_plain_geomean"] = float(np.exp(np.mean(np.log(sp))))
        summary[k]["eig_speedup_vs_plain_geomean"] = float(np.exp(np.mean(np.log(se))))
        ang = [r[k]["angle_vs_plain"] for r in per_image if "angle_vs_plain" in r[k]]
        if ang:
            summary[k]["angle_vs_plain_median"] = float(np.median(ang))
            summary[k]["angle_vs_plain_max"] = float(np.max(ang))
    out = dict(workload=f"{args.count} images {S}x{S}, three-class Voronoi, {args.regions} regions, sigma 8, seeds "
                        f"{args.seed0}..{args.seed0 + args.count - 1}, 9 eigenvectors, residual {args.tol:g}",
               block=dict(basis=args.basis, keep=args.keep), plain=dict(basis=args.plain_basis, keep=args.plain_keep), summary=summary, images=per_image)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

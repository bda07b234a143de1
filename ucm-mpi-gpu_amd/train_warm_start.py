"""Train the warm-start network on eigenvectors the method computes itself.

Training set: n seeded synthetic images (default 2000 images of 512 x 512,
seeds 1000.., 50-300 Voronoi regions drawn per seed, sigma 8, three region
classes), each with its own texton codebook (seeded k-means++, 10 iterations,
on 8192 sampled Schmid responses).  For every image the GPU runs the path to
the eigenvectors (plain Lanczos, residual 1e-4) and keeps the cue features
and the D-weighted 4x4-cell sums of eigenvectors 1..8 on the device; the
last --val images are held out.  Adam trains the network in phases of
decreasing learning rate.  Every --select-every epochs the current parameters
start block Lanczos on separate validation images at the target scale
(default two 4096 x 4096 images, 1600 regions, seeds 6000 and 6001, disjoint
from the evaluation seeds); the parameters with the fewest operator
applications there (ties: lower held-out loss) are the ones written, as raw
little-endian FP32 (the layout of gpb_warm_get_params) with a JSON record
beside them.

    python ucm-mpi-gpu_amd/train_warm_start.py --out gpurun_out/warm_start.f32
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
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


def regions_for(seed: int, lo: int = 50, hi: int = 300) -> int:
    """Region count of a training image, a fixed function of its seed (SplitMix64)."""
    z = (seed * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) & (2 ** 64 - 1)
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
    z ^= z >> 31
    return lo + int(z % (hi - lo + 1))


def make_image(gpb, size: int, seed: int, samples: int):
    rgb = gpb.synth_voronoi(size, size, regions_for(seed), 8.0, seed, classes=3)
    packed = gpb.pack_rgb(rgb)
    return packed, gpb.make_codebook(packed, seed=seed, iters=10, max_samples=samples)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2000, help="images (the last --val are held out)")
    ap.add_argument("--val", type=int, default=100)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--samples", type=int, default=8192, help="Schmid responses sampled per codebook")
    ap.add_argument("--tol", type=float, default=1e-4, help="residual target of the training eigenvectors")
    ap.add_argument("--phases", default="16:1e-3,12:3e-4,8:1e-4", help="epochs:lr,... trained in order")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--workers", type=int, default=8, help="threads building images and codebooks")
    ap.add_argument("--augment", type=int, default=1, help="random symmetries of the square cell grid (0: off)")
    ap.add_argument("--select-size", type=int, default=4096)
    ap.add_argument("--select-regions", type=int, default=1600)
    ap.add_argument("--select-seeds", default="6000,6001")
    ap.add_argument("--select-every", type=int, default=3, help="epochs between checkpoint evaluations (0: off)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    gpb = _binding()
    ctx = gpb.Context(0)
    ctx.set_eig_options(nvec=9, basis=40, keep=20, tol=args.tol, max_matvecs=40000)
    ctx.warm_init(args.init_seed)
    ctx.warm_dataset_reset(args.size, args.size, args.n)
    t0 = time.perf_counter()
    seeds = [args.seed0 + i for i in range(args.n)]
    with cf.ThreadPoolExecutor(args.workers) as ex:
        futs = [ex.submit(make_image, gpb, args.size, s, args.samples) for s in seeds]
        for i, f in enumerate(futs):
            packed, cb = f.result()
            ctx.warm_dataset_add(packed, cb)
            futs[i] = None
            if (i + 1) % 100 == 0:
                print(f"training set: {i + 1}/{args.n} images ({time.perf_counter() - t0:.1f} s)", flush=True)
    t_data = time.perf_counter() - t0
    n_train = args.n - args.val
    rec = dict(images=args.n, held_out=args.val, size=args.size, seeds=[args.seed0, args.seed0 + args.n - 1],
               regions="50-300 (per seed)", classes=3, sigma=8.0, codebook_samples=args.samples,
               eig_tol=args.tol, batch=args.batch, init_seed=args.init_seed, augment=bool(args.augment),
               data_s=round(t_data, 1))
    val0 = ctx.warm_eval_loss(n_train, args.val)
    print(f"held-out loss before training {val0:.4f}", flush=True)
    rec["val_loss_init"] = val0
    rec["phases"] = []
    # target-scale selection images (a second context runs the solves, so the
    # training set's buffers stay untouched)
    sel = []
    if args.select_every > 0:
        for s in [int(x) for x in args.select_seeds.split(",") if x]:
            rgb = gpb.synth_voronoi(args.select_size, args.select_size, args.select_regions, 8.0, s, classes=3)
            p = gpb.pack_rgb(rgb)
            sel.append((p, gpb.make_codebook(p, seed=s, iters=10)))
        sctx = gpb.Context(0)
        sctx.set_eig_options(nvec=9, basis=40, keep=20, tol=1e-3, max_matvecs=40000)

    def select_score():
        sctx.set_warm_params(ctx.warm_params())
        sctx.set_warm_start(1, 64, 24)
        mv = []
        for p, cb in sel:
            sctx.compute(p, cb)
            st = sctx.stats()
            mv.append(st["matvecs"] + (100000 if st["fallback"] else 0))
        return float(np.mean(mv))

    best = None
    rec["selection"] = []
    t1 = time.perf_counter()
    epoch = 0
    for k, ph in enumerate(args.phases.split(",")):
        ep, lr = ph.split(":")
        ep, lr = int(ep), float(lr)
        losses = []
        for e in range(ep):
            losses.append(float(ctx.warm_train(0, n_train, 1, args.batch, lr, 7919 * (k + 1) + e,
                                               bool(args.augment))[0]))
            epoch += 1
            print(f"phase {k} (lr {lr:g}) epoch {e + 1}/{ep}: training loss {losses[-1]:.4f}", flush=True)
            if sel and epoch % args.select_every == 0:
                score = select_score()
                val = ctx.warm_eval_loss(n_train, args.val)
                rec["selection"].append(dict(epoch=epoch, matvecs=score, val_loss=val))
                print(f"  epoch {epoch}: {args.select_size}^2 validation matvecs {score:.0f}, held-out loss "
                      f"{val:.4f}", flush=True)
                if best is None or (score, val) < (best[0], best[1]):
                    best = (score, val, epoch, ctx.warm_params())
        val = ctx.warm_eval_loss(n_train, args.val)
        print(f"phase {k}: held-out loss {val:.4f}", flush=True)
        rec["phases"].append(dict(epochs=ep, lr=lr, train_loss=losses, val_loss=val))
    rec["train_s"] = round(time.perf_counter() - t1, 1)
    params = ctx.warm_params()
    if best is not None:
        params = best[3]
        rec["selected"] = dict(epoch=best[2], matvecs=best[0], val_loss=best[1],
                               images=f"{args.select_size}x{args.select_size}, {args.select_regions} regions, seeds "
                                      f"{args.select_seeds}")
        sctx.close()
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    params.astype("<f4").tofile(out)
    rec["param_count"] = int(params.size)
    out.with_suffix(".json").write_text(json.dumps(rec, indent=1))
    ctx.close()
    print(json.dumps(dict(val_loss_init=val0, val_loss=rec["phases"][-1]["val_loss"], data_s=rec["data_s"],
                          train_s=rec["train_s"], selected=rec.get("selected", {}).get("epoch"))), flush=True)


if __name__ == "__main__":
    main()

"""gPb throughput benchmark on one MI355X.

Workload: plain gPb of one 2048x2048 seeded Voronoi image (400 regions,
sinusoid textures, noise sigma 8, seed 2), 9 eigenvectors to a relative
generalised residual <= 1e-3, FP32 with FP64 Lanczos reorthogonalisation,
Ritz recombination and projected eigensolve.  A step is one whole gPb pass
(texton -> cues -> nonmax -> W -> eigensolve -> sPb -> gPb) with the packed
image and codebook already resident in HBM and the outputs left in HBM.

Prints one JSON line (rank 0).  One untimed step with HIP events around every
launch gives `kernels` (device ms, launches and GB/s per kernel class for that
step) and names the class with the largest device time; in the timed steps
events bracket only that class's launches on the library's stream, and
`roofline` describes it from those; its `copy_gbs` is the rate (bytes read +
written) of a plain 1 GiB device copy on the same GPU, the practical streaming
ceiling next to the published peak.

`cpu_baseline` times the reference's own CPU path (oracle/ref_cpu_path.py: its
lib_image cues, buildW affinity and Spectra eigensolver, compiled from the
reference tree into oracle/_ref, with the oracle's port for the texton and
sPb stages the reference cannot build here) on 256x256 patches of the same
image -- the reference's deployment shape, one process per patch (an 8x8
Img2Pieces cut of the 2048x2048 image) -- on the host's cores: one patch
alone on one thread, and up to 16 patches at once, one process each.  It runs
first, in child processes, before this process touches the GPU.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def copy_rate_gbs(torch, mib: int = 1024, reps: int = 10) -> float:
    """Bytes read + written per second of a plain device-to-device copy of mib MiB."""
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    rate = 2.0 * n * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return rate


def cpu_gen(path: str, size: int, regions: int, patch: int, count: int):
    """Child mode: the bench image's first `count` patch x patch tiles (row-major
    over the image) and its codebook, written to an .npz for the CPU run."""
    import numpy as np

    gpb = _load("gpb_amd_binding", ROOT / "ucm-mpi-gpu_amd" / "gpb.py")
    packed = gpb.pack_rgb(gpb.synth_voronoi(size, size, regions, 8.0, 2))
    cb = gpb.make_codebook(packed, seed=2, iters=10)
    g = size // patch
    tiles = [packed[(i // g) * patch:(i // g + 1) * patch, (i % g) * patch:(i % g + 1) * patch]
             for i in range(min(count, g * g))]
    np.savez(path, patches=np.stack(tiles), codebook=cb)


def cpu_baseline(size: int, regions: int, patch: int = 256, workers: int = 16):
    """The reference's own CPU path on patch x patch tiles of the bench image,
    run in child processes (see the module docstring); None when oracle/_ref is
    not built."""
    import subprocess
    import tempfile

    if not all((ROOT / "oracle" / "_ref" / n).exists()
               for n in ("libref_libimage.so", "libref_buildw.so", "libref_spectra.so")):
        return None
    workers = max(1, min(workers, os.cpu_count() or 1, len(os.sched_getaffinity(0))))
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    with tempfile.TemporaryDirectory() as td:
        npz, res = os.path.join(td, "patches.npz"), os.path.join(td, "cpu.json")
        # the children's output (buildW prints progress bars) stays off this process' stdout
        for cmd in ([sys.executable, str(ROOT / "bench.py"), "--cpu-gen", npz, "--size", str(size), "--regions",
                     str(regions), "--cpu-patch", str(patch), "--cpu-workers", str(workers)],
                    [sys.executable, str(ROOT / "oracle" / "ref_cpu_path.py"), "--input", npz, "--workers",
                     str(workers), "--out", res]):
            p = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if p.returncode != 0:
                raise RuntimeError(f"CPU baseline step failed ({p.returncode}): {' '.join(cmd)}\n{p.stdout[-2000:]}")
        r = json.loads(Path(res).read_text())
    px = patch * patch
    mean_patch = sum(r["parallel_patch_seconds"]) / len(r["parallel_patch_seconds"])
    n_c2 = (size // patch) ** 2
    return dict(value=round(r["workers"] * px / r["parallel_wall_seconds"] / 1e6, 6), unit="Mpixel/s",
                cores=r["workers"], kind="reference",
                sample=(f"{r['workers']} of the {n_c2} {patch}x{patch} patches of the {size}x{size} image, one process "
                        f"(one thread) per patch as the reference runs one MPI rank per patch: the reference's own "
                        f"lib_image cues, buildW and Spectra SymGEigsSolver (nev 9, ncv 18, tol 1e-10) compiled "
                        f"from its tree, the oracle's port for textons and sPb/gPb; "
                        f"{r['parallel_wall_seconds']:.1f} s wall"),
                single_thread=dict(value=round(px / r["single_seconds"] / 1e6, 6), seconds=round(r["single_seconds"], 2),
                                   stage_seconds={k: round(v, 2) for k, v in r["single_stages"].items()},
                                   spectra_operations=r["single_operations"], converged=r["single_converged"]),
                mean_patch_seconds=round(mean_patch, 2), host_nproc=r["nproc"], host_affinity=r["affinity"],
                whole_image_seconds_est=round(n_c2 * mean_patch / r["workers"], 1),
                c3_4096_patch_split_seconds_est=round((4096 // patch) ** 2 * mean_patch / r["workers"], 1))


def measured_traffic(kernel_class: str, size: int):
    """HBM bytes per launch of a kernel class, from the committed rocprofv3 PMC
    summary (FETCH_SIZE x its copy-kernel calibration + WRITE_SIZE, per
    dispatch, averaged over the class's dispatches) for this image size, or
    None when no summary covers it."""
    path = ROOT / "profiles" / f"traffic_{size}.json"
    if not path.exists():
        return None
    d = json.loads(path.read_text())
    c = d.get("classes", {}).get(kernel_class)
    return round(c["bytes_per_launch"]) if c else None


def _form_roofline(ctx, run):
    """One more untimed run with HIP events around every launch: the kernel
    class with the largest device time, its algorithmic bytes over its event
    time, against the 8 TB/s HBM peak."""
    import torch

    ctx.set_profiling(True)
    run()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    dom = max(kt, key=lambda k: kt[k]["ms"])
    d = kt[dom]
    total = sum(v["ms"] for v in kt.values())
    gbs = d["bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else 0.0
    return dict(kernel=dom, bound="hbm", achieved=round(gbs, 1), peak=8000.0, unit="GB/s", frac=round(gbs / 8000.0, 4),
                device_ms=round(d["ms"], 2), launches=d["launches"], share_of_device_time=round(d["ms"] / total, 3))


def bench_patches(gpb, ctx, tol: float, size: int = 5000, regions: int = 2500, grid: int = 4, overlap: int = 16,
                  seed: int = 4):
    """The batched form: a 5000x5000 seeded Voronoi image (2500 regions, sigma 8,
    seed 4) cut into a 4x4 grid with seeded cut offsets and a 16-px overlap, the
    16 patches' whole gPb computed in one batched Lanczos (mode 1) and, for
    comparison, with one solve per patch (mode 0); stitched outputs in HBM.
    Mpixel/s counts the image's pixels once (not the overlap)."""
    import numpy as np
    import torch

    rgb = gpb.synth_voronoi(size, size, regions, 8.0, seed)
    packed = gpb.pack_rgb(rgb)
    del rgb
    cb = gpb.make_codebook(packed, seed=seed, iters=10)
    rects, cores = gpb.cut_patches(size, size, grid, overlap, seed)
    dev = torch.device("cuda:0")
    d_rgb = torch.from_numpy(packed.view(np.int32)).to(dev)
    d_cb = torch.from_numpy(cb).to(dev)
    d_gmax = torch.zeros((size, size), dtype=torch.float32, device=dev)
    d_orient = torch.zeros((8, size, size), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    out = {}
    for name, mode in (("batched", 1), ("one_by_one", 0)):
        run = lambda: ctx.compute_patches_device(size, size, d_rgb.data_ptr(), d_cb.data_ptr(), rects, cores, mode,
                                                 d_gmax.data_ptr(), d_orient.data_ptr())
        run()  # warm-up (tile and workgroup tuning, workspace)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = [ctx.patch_stats(p) for p in range(len(rects))]
        out[name] = dict(value=round(size * size / dt / 1e6, 4), unit="Mpixel/s", ms=round(dt * 1e3, 1),
                         matvecs_total=int(sum(x["matvecs"] for x in st)),
                         max_residual=max(x["max_residual"] for x in st),
                         stage_ms=[round(x, 1) for x in st[0]["stage_ms"]])
        if mode == 1:
            out[name]["roofline"] = _form_roofline(ctx, run)
    out["speedup_batched_vs_one_by_one"] = round(out["one_by_one"]["ms"] / out["batched"]["ms"], 3)
    out["workload"] = (f"{size}x{size} seeded Voronoi ({regions} regions, sigma 8, seed {seed}), {grid}x{grid} patches "
                       f"{int(rects[:, 2].min())}..{int(rects[:, 2].max())} x {int(rects[:, 3].min())}.."
                       f"{int(rects[:, 3].max())} px with {overlap}-px overlap, 9 eigenvectors each, residual <= {tol:g}")
    del d_rgb, d_gmax, d_orient
    torch.cuda.empty_cache()
    return out


def bench_warm(gpb, ctx, step, S: int, steps: int, tol: float):
    """The warm-start form on the same image: block Lanczos (8 columns plus the
    locked D^1/2 1) started from the trained network's block (mode 1), beside
    plain Lanczos (mode 0) and block Lanczos from a random block (mode 2), all
    timed without per-kernel events, at the same residual target."""
    import numpy as np
    import torch

    weights = ROOT / "ucm-mpi-gpu_amd" / "weights" / "warm_start.f32"
    if not weights.exists():
        return dict(skipped=f"no trained parameters at {weights.relative_to(ROOT)}")
    ctx.set_warm_params(np.fromfile(weights, dtype="<f4"))
    out = {}
    for name, mode in (("plain", 0), ("warm_start", 1), ("random_block", 2)):
        ctx.set_warm_start(mode, 80, 24)
        step()  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = ctx.stats()
        out[name] = dict(value=round(S * S * steps / dt / 1e6, 4), unit="Mpixel/s", ms=round(dt * 1e3 / steps, 1),
                         matvecs=st["matvecs"], fallback=st["fallback"], max_residual=st["max_residual"],
                         network_ms=round(st["warm_ms"], 2), stage_ms=[round(x, 1) for x in st["stage_ms"]])
        if mode == 1:
            out[name]["roofline"] = _form_roofline(ctx, step)
    ctx.set_warm_start(0)
    out["speedup_warm_vs_plain"] = round(out["plain"]["ms"] / out["warm_start"]["ms"], 3)
    out["speedup_random_block_vs_plain"] = round(out["plain"]["ms"] / out["random_block"]["ms"], 3)
    out["workload"] = f"the plain form's image, 9 eigenvectors, residual <= {tol:g} (measured), block basis 80 keep 24"
    return out


def bench_parity(gpb, ctx, step, S: int, tol_rel: float = 3e-4):
    """The plain form at the parity setting: convergence when every pair's
    residual is within tol_rel times its Ritz value (the reference CPU solver's
    rule, Spectra num_converged), where the eigenvalue (1e-5 relative),
    subspace (1e-3 rad) and gPb (1e-3) bars hold against the all-FP64 solve on
    this image (tests/test_configs_scale.py); one timed step."""
    import torch

    ctx.set_eig_convergence(True)
    ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=tol_rel, max_matvecs=80000)
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = ctx.stats()
    finally:
        ctx.set_eig_convergence(False)
        ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=1e-3, max_matvecs=40000)
    return dict(value=round(S * S / dt / 1e6, 4), unit="Mpixel/s", ms=round(dt * 1e3, 1), matvecs=st["matvecs"],
                restarts=st["restarts"], stage_ms=[round(x, 1) for x in st["stage_ms"]],
                workload=f"the plain form's image, 9 eigenvectors, residual <= {tol_rel:g} x Ritz value (basis 48, "
                         f"keep 16), FP64 Rayleigh-Ritz refinement")


def bench_host_to_host(gpb, ctx, packed, codebook, steps: int):
    """The plain form through the host-buffer entry point: upload of the packed
    image and codebook, the whole path, download of the gPb maximum and the 8
    oriented maps (computeGPb's outputs)."""
    gpb_out = None
    ctx.compute(packed, codebook, orient=False)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        gpb_out = ctx.compute(packed, codebook, orient=False)
    dt = (time.perf_counter() - t0) / steps
    S = packed.shape[0]
    del gpb_out
    return dict(value=round(S * S / dt / 1e6, 4), unit="Mpixel/s", ms=round(dt * 1e3, 1),
                workload="the plain form's image, host buffers in and out (4 B/px up; 4 + 32 B/px of gPb maps and "
                         "the mPb_nmax map down)")


def bench_ucm(gpb, ctx, S: int, d_rgb, d_cb, gpb_ms: float):
    """The consumer of the gPb maps on the plain form's image: gpb_compute_ucm
    (gPb on the GPU, then OWT-UCM: the per-pixel stages on the GPU, the
    watershed, contour fitting, cleaning and merging on the host), beside the
    CPU oracle of the consumer (a Python restatement of the reference's
    steps; the reference's own Contours2Ucm needs OpenCV) and the host-only
    consumer on a 160 x 160 crop of the same image's oriented maps."""
    import importlib.util
    import numpy as np
    import torch

    ctx.compute_ucm(S, S, d_rgb.data_ptr(), d_cb.data_ptr(), True)  # warm-up
    ctx.set_profiling_classes(["ucm"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ucm = ctx.compute_ucm(S, S, d_rgb.data_ptr(), d_cb.data_ptr(), True)
    dt = time.perf_counter() - t0
    kt = ctx.kernel_times()["ucm"]
    ctx.set_profiling_classes([])
    out = dict(value=round(S * S / dt / 1e6, 4), unit="Mpixel/s", ms=round(dt * 1e3, 1),
               ucm_ms=round(dt * 1e3 - gpb_ms, 1), gpb_ms=round(gpb_ms, 1),
               ucm_kernels=dict(ms=round(kt["ms"], 3), launches=kt["launches"],
                                gbs=round(kt["bytes"] / (kt["ms"] * 1e-3) / 1e9, 1) if kt["ms"] > 0 else 0.0),
               ucm_nonzero=int(np.count_nonzero(ucm)),
               workload=f"the plain form's image: gPb, then the ultrametric contour map ({2 * S + 1} x {2 * S + 1})")
    # the oracle and the host consumer on a crop of the same maps
    spec = importlib.util.spec_from_file_location("ucm_oracle_bench", ROOT / "oracle" / "ucm_oracle.py")
    uo = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(uo)
    packed = np.empty((S, S), np.uint32)
    d_rgb_host = d_rgb.cpu().numpy().view(np.uint32)
    packed[:] = d_rgb_host.reshape(S, S)
    crop = 160
    c0 = S // 2 - crop // 2
    orient = ctx.compute(packed, d_cb.cpu().numpy())["gpb_orient"].astype(np.float64)
    maps = np.ascontiguousarray(orient[:, c0:c0 + crop, c0:c0 + crop])
    t0 = time.perf_counter()
    gpb.contours2ucm(maps, True)
    t_host = time.perf_counter() - t0
    t0 = time.perf_counter()
    uo.contours2ucm(maps, True)
    t_orc = time.perf_counter() - t0
    out["cpu_sample"] = dict(sample=f"{crop}x{crop} crop of the image's oriented gPb maps", oracle_seconds=round(t_orc, 3),
                             host_consumer_seconds=round(t_host, 4), kind="port",
                             oracle_mpixel_s=round(crop * crop / t_orc / 1e6, 5))
    # the host stages of the consumer on the whole image's maps
    st = gpb.ucm_stages(orient)
    out["host_stage_ms"] = st["stage_ms"]
    out["regions"] = st["regions"]
    dump = os.environ.get("GPB_UCM_DUMP_MAPS")
    if dump:  # a 1024 x 1024 crop of the oriented maps, for profiling the host stages elsewhere
        q0 = S // 2 - 512
        np.ascontiguousarray(orient[:, q0:q0 + 1024, q0:q0 + 1024], np.float32).tofile(dump)
    return out


def bench_c3(gpb, ctx, tol: float, cpu):
    """The 4096x4096 three-class image (1600 regions, sigma 8, seed 3): the plain
    form's whole path, device-resident, beside the reference CPU path's time for
    the same image in its deployment shape (16 x 16 patches of 256 x 256, one
    process per patch on the host's cores, estimated from the patches the CPU
    baseline timed)."""
    import numpy as np
    import torch

    S = 4096
    packed = gpb.pack_rgb(gpb.synth_voronoi(S, S, 1600, 8.0, 3, 3))
    cb = gpb.make_codebook(packed, seed=3, iters=10)
    dev = torch.device("cuda:0")
    d_rgb = torch.from_numpy(packed.view(np.int32)).to(dev)
    d_cb = torch.from_numpy(cb).to(dev)
    d_gmax = torch.empty((S, S), dtype=torch.float32, device=dev)
    d_gall = torch.empty((8, S, S), dtype=torch.float32, device=dev)
    run = lambda: ctx.compute_device(S, S, d_rgb.data_ptr(), d_cb.data_ptr(), d_gmax.data_ptr(), d_gall.data_ptr())
    run()  # warm-up (tile tuning for the shape, workspace)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = ctx.stats()
    out = dict(value=round(S * S / dt / 1e6, 4), unit="Mpixel/s", ms=round(dt * 1e3, 1), matvecs=st["matvecs"],
               stage_ms=[round(x, 1) for x in st["stage_ms"]],
               workload=f"4096x4096 three-class Voronoi (1600 regions, sigma 8, seed 3), 9 eigenvectors, residual "
                        f"<= {tol:g}")
    if cpu is not None:
        est = cpu["c3_4096_patch_split_seconds_est"]
        out["cpu_reference_seconds_est"] = est
        out["speedup_vs_cpu_reference"] = round(est / dt, 1)
        out["cpu_cores"] = cpu["cores"]
    del d_rgb, d_gmax, d_gall
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--regions", type=int, default=400)
    ap.add_argument("--cpu-patch", type=int, default=256)
    ap.add_argument("--cpu-workers", type=int, default=16)
    ap.add_argument("--cpu-gen", default="", help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=1, help="Lanczos block size (1 = single-vector)")
    ap.add_argument("--basis", type=int, default=0)
    ap.add_argument("--keep", type=int, default=0)
    ap.add_argument("--spmv-tile", type=int, default=0, help="SpMV tile code 10 * wavefronts + rows per thread (42, 44, 48, 82, 84, 161, 162); 0 = autotune")
    ap.add_argument("--spmv-fp32", action="store_true", help="SpMV arithmetic in FP32 (default FP64 arithmetic)")
    ap.add_argument("--full-reorth", action="store_true",
                    help="whole-basis Gram-Schmidt on every Lanczos step (default: partial reorthogonalisation)")
    ap.add_argument("--no-kernel-timing", action="store_true", help="time the steps without per-kernel HIP events")
    ap.add_argument("--forms", default="parity,warm,batched,host,ucm,c3",
                    help="comma list of extra forms to time: parity (the Spectra-rule setting), warm (learned warm "
                         "start), batched (C4), host (host buffers in and out), ucm (gPb then OWT-UCM), c3 "
                         "(4096x4096), none")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.cpu_gen:
        cpu_gen(args.cpu_gen, args.size, args.regions, args.cpu_patch, args.cpu_workers)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the gPb path runs one image (or one patch batch) per GPU, and this "
              "build has no multi-GPU launcher; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline first, in child processes, before this process touches the GPU
    cpu = None if (args.no_cpu or args.copy_probe) else cpu_baseline(args.size, args.regions, args.cpu_patch,
                                                                     args.cpu_workers)

    import numpy as np
    import torch

    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_probe_bytes_read=n * 4, copy_probe_bytes_written=n * 4, launches=3)), flush=True)
        return

    gpb = _load("gpb_amd_binding", ROOT / "ucm-mpi-gpu_amd" / "gpb.py")
    S, tol = args.size, 1e-3
    rgb = gpb.synth_voronoi(S, S, args.regions, 8.0, 2)
    packed = gpb.pack_rgb(rgb)
    codebook = gpb.make_codebook(packed, seed=2, iters=10)
    dev = torch.device("cuda:0")
    d_rgb = torch.from_numpy(packed.view(np.int32)).to(dev)
    d_cb = torch.from_numpy(codebook).to(dev)
    d_gmax = torch.empty((S, S), dtype=torch.float32, device=dev)
    d_gall = torch.empty((8, S, S), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    ctx = gpb.Context(0)
    basis = args.basis or (48 if args.block == 1 else 192)
    keep = args.keep or (16 if args.block == 1 else 64)
    ctx.set_eig_options(nvec=9, basis=basis, keep=keep, tol=tol, max_matvecs=40000)
    ctx.set_block_size(args.block)
    ctx.set_spmv_tile(args.spmv_tile)
    ctx.set_spmv_fp64(not args.spmv_fp32)
    ctx.set_reorth(not args.full_reorth)

    def step():
        ctx.compute_device(S, S, d_rgb.data_ptr(), d_cb.data_ptr(), d_gmax.data_ptr(), d_gall.data_ptr())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # One untimed step with events on every kernel class gives the per-class
    # breakdown and names the dominant kernel; the timed steps then carry
    # events on that class only, so the other launches run as without them.
    dom = None
    if not args.no_kernel_timing:
        ctx.set_profiling(True)
        step()
        breakdown = ctx.kernel_times()
        ctx.set_profiling(False)
        dom = max(breakdown, key=lambda k: breakdown[k]["ms"])
        ctx.set_profiling_classes([dom])
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        stats.append(ctx.stats())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt * 1e3 / args.steps
    value = S * S * args.steps / dt / 1e6

    if args.no_kernel_timing:
        st = stats[-1]
        print(json.dumps(dict(value=round(value, 4), ms_per_step=round(ms_step, 2), kernel_timing=False,
                              lanczos_basis=st["basis"], lanczos_matvecs=st["matvecs"],
                              lanczos_restarts=st["restarts"], max_residual=st["max_residual"],
                              stage_ms=[round(x, 2) for x in st["stage_ms"]])), flush=True)
        return
    d = ctx.kernel_times()[dom]
    ctx.set_profiling(False)
    avg_ms = d["ms"] / max(1, d["launches"])
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else 0.0
    traffic = measured_traffic(dom, S)
    copy_gbs = copy_rate_gbs(torch)
    roofline = dict(kernel=dom, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    copy_gbs=round(copy_gbs, 1), frac_of_copy=round(achieved / copy_gbs, 4),
                    avg_launch_us=round(avg_ms * 1e3, 2), launches=d["launches"],
                    bytes_per_launch=round(d["bytes"] / max(1, d["launches"])),
                    share_of_step=round(d["ms"] / (ms_step * args.steps), 4))
    st = stats[-1]
    line = dict(metric="gPb Mpixel/s on 1 MI355X at eig-residual <=1e-3 (plain Lanczos)", value=round(value, 4),
                unit="Mpixel/s", n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(ms_step, 2),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32+f64",
                data="synthetic",
                config=dict(workload=f"plain gPb, {S}x{S} seeded Voronoi ({args.regions} regions, sigma 8, seed 2), "
                                     f"9 eigenvectors, residual <= {tol:g}",
                            form="plain", lanczos_block=st["block"], lanczos_basis=st["basis"],
                            lanczos_matvecs=st["matvecs"],
                            lanczos_restarts=st["restarts"], reorth_second_passes=st["second_passes"],
                            max_residual=st["max_residual"], spmv_tile=ctx.spmv_tile(S, S), spmv_arith="f32" if args.spmv_fp32 else "f64",
                            stage_ms=[round(x, 2) for x in st["stage_ms"]]),
                roofline=roofline,
                kernels={k: dict(ms=round(v["ms"], 3), launches=v["launches"],
                                 gbs=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else 0.0)
                         for k, v in breakdown.items()})
    forms = [f for f in args.forms.split(",") if f and f != "none"]
    if forms:
        line["forms"] = {}
    if "parity" in forms:
        line["forms"]["parity_setting"] = bench_parity(gpb, ctx, step, S)
    if "warm" in forms:
        line["forms"]["warm_start"] = bench_warm(gpb, ctx, step, S, args.steps, tol)
    if "host" in forms:
        line["forms"]["host_to_host"] = bench_host_to_host(gpb, ctx, packed, codebook, args.steps)
    if "batched" in forms:
        line["forms"]["batched_patches"] = bench_patches(gpb, ctx, tol)
    if "ucm" in forms:
        line["forms"]["ucm"] = bench_ucm(gpb, ctx, S, d_rgb, d_cb, ms_step)
    if "c3" in forms:
        line["forms"]["c3_4096"] = bench_c3(gpb, ctx, tol, cpu)
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

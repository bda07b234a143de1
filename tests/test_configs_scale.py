"""Parity at the configurations' own sizes, on the GPU.

C2 (2048x2048, 400 regions, sigma 8, seed 2):
  * front end: textons bit-exact, mPb_all and mPb_nmax within 1e-4 and W within
    1e-5 relative of the CPU oracle, on four 64x64 windows at random interior
    positions; the oracle runs on each window grown by a 56-pixel margin (the
    Schmid support 24, the largest half-disc 20, nonmax 1 and the stencil
    radius 5 fit inside it, so the window's values equal the whole image's);
  * eigenpairs and gPb at the parity setting -- residual <= 3e-4 times the Ritz
    value, the reference CPU solver's rule (Spectra num_converged) -- against
    the all-FP64 GPU eigensolve on the same mPb_nmax (that solver is pinned to
    the oracle's LAPACK-pinned eigensolver in test_precision.py): eigenvalues
    within 1e-5 relative, the eigenvector subspace within 1e-3 rad, and gPb
    within 1e-3 (w_sPb times the largest sPb difference, the only term of the
    oriented gPb the eigensolve changes).
C3 (a 1024x1024 window of the 4096x4096 three-class image, seed 3): the
  production path at the parity setting against the all-FP64 GPU path: mPb_nmax
  within 1e-4 off the few pixels whose nonmax survival flips, and the
  eigenpairs and gPb to the bars above (FP64 solve on the production mPb, which
  isolates the precision of the solver from the cue stage).
C4 (5000x5000, 2500 regions, seed 4, a 4x4 cut with a 16-pixel overlap, one
  batched Lanczos at residual 1e-3): every patch's nine returned pairs have
  their residual recomputed in FP64 on the host from the patch's own W and
  vectors (the oracle's stencil code) within the target; a window of each of
  patches 0, 5, 10 and 15 against the oracle's front end as for C2.
C5 (an unseen 4096x4096 image of the evaluation set, seed 5000): the shipped
  network's start needs fewer operator applications than plain Lanczos and
  than the same block solver from a random start, at the bench's residual
  target, with no fallback and every measured residual within it.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import scipy.linalg as sl

from conftest import ROOT

W_SPB = 0.0074  # the sPb weight of the gPb sum (src/globalPb.cpp:31-37)
PARITY_TOL = 3e-4
MARGIN = 56
UPPER = [(dy, dx) for dy in range(6) for dx in range(-5, 6) if dx * dx + dy * dy <= 25 and (dy > 0 or dx > 0)]


def _angle(Va, Vb, D):
    s = np.sqrt(D.reshape(-1).astype(np.float64))
    A = (Va.reshape(Va.shape[0], -1).astype(np.float64) * s).T
    B = (Vb.reshape(Vb.shape[0], -1).astype(np.float64) * s).T
    return float(np.max(sl.subspace_angles(A, B)))


def _image(gpb, S, regions, seed, classes=1):
    rgb = gpb.synth_voronoi(S, S, regions, 8.0, seed, classes)
    packed = gpb.pack_rgb(rgb)
    return packed, gpb.make_codebook(packed, seed=seed, iters=10)


def _windows(rng, H, W, n, size=64):
    lo = MARGIN
    return [(int(rng.integers(lo, H - lo - size)), int(rng.integers(lo, W - lo - size))) for _ in range(n)]


def _front_end_window(orc, ctx, packed, cb, t_gpu, mpb_gpu, nm_gpu, nbins, y, x, size=64):
    """Oracle front end on the window grown by MARGIN, with the whole image's
    histogram lengths (nbins: L, a, b, textons -- hist_gradient_2D sizes its
    histograms by the largest label of its input, and the bin smoothing sees
    that length); GPU values of the whole image on the window; W from the GPU
    on the grown crop of its own mPb_nmax."""
    ys, xs = slice(y - MARGIN, y + size + MARGIN), slice(x - MARGIN, x + size + MARGIN)
    crop = np.ascontiguousarray(packed[ys, xs])
    t = orc.texton(crop, cb)
    _, mpb = orc.local_cues(crop, t, nbins)
    nm = orc.nonmax(mpb)
    win = (slice(MARGIN, MARGIN + size), slice(MARGIN, MARGIN + size))
    gw = (slice(y, y + size), slice(x, x + size))
    assert np.array_equal(t[win], t_gpu[gw])
    err_mpb = np.abs(mpb[(slice(None),) + win] - mpb_gpu[(slice(None),) + gw]).max()
    flips = (nm[win] > 0) != (nm_gpu[gw] > 0)
    err_nm = np.abs(nm[win] - nm_gpu[gw])[~flips].max()
    assert err_mpb <= 1e-4 and err_nm <= 1e-4 and np.count_nonzero(flips) <= 2, (err_mpb, err_nm, flips.sum())
    # W on the window, where no nonmax flip lies within the stencil radius
    # (flips counted on the window grown by that radius)
    Wd, _ = orc.build_w(nm)
    wu, _ = ctx.build_w(np.ascontiguousarray(nm_gpu[ys, xs]))
    upper_idx = [k for k, (dx, dy) in enumerate(orc.DISC) if dy > 0 or (dy == 0 and dx > 0)]
    g5 = (slice(MARGIN - 5, MARGIN + size + 5), slice(MARGIN - 5, MARGIN + size + 5))
    f5 = (nm[g5] > 0) != (nm_gpu[ys, xs][g5] > 0)
    near = np.zeros((size, size), bool)
    for dy in range(11):
        for dx in range(11):
            near |= f5[dy:dy + size, dx:dx + size]
    a, b = Wd[upper_idx][(slice(None),) + win], wu[(slice(None),) + win]
    rel = (np.abs(a - b) / np.maximum(np.abs(a), 1e-30))[:, ~near]
    assert rel.max() <= 1e-5, rel.max()
    return err_mpb, err_nm, int(flips.sum())


def _nbins(ctx, packed, t):
    _, nb = ctx.lab_quantize(packed)
    return [int(nb[0]), int(nb[1]), int(nb[2]), int(t.max()) + 1]


def _full_disc(orc, wup):
    """The oracle's 81-plane stencil (unit centre) from the 40 upper planes."""
    _, H, W = wup.shape
    full = np.zeros((len(orc.DISC), H, W))
    for k, (dx, dy) in enumerate(orc.DISC):
        if dx == 0 and dy == 0:
            full[k] = 1.0
        elif (dy, dx) in UPPER:
            full[k] = wup[UPPER.index((dy, dx))]
        else:
            u = UPPER.index((-dy, -dx))  # W(p, p + e) = W(p + e, p) = upper plane at p + e
            ys, yd = (slice(0, H + dy), slice(-dy, H)) if dy < 0 else (slice(dy, H), slice(0, H - dy))
            xs, xd = (slice(0, W + dx), slice(-dx, W)) if dx < 0 else (slice(dx, W), slice(0, W - dx))
            full[k][yd, xd] = wup[u][ys, xs]
    return full


# ---- C2 ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c2(gpb, ctx):
    print("C2 2048x2048: image, codebook, GPU front end", flush=True)
    packed, cb = _image(gpb, 2048, 400, 2)
    t = ctx.texton(packed, cb)
    _, mpb = ctx.local_cues(packed, t, with_cues=False)
    nm = ctx.nonmax(mpb)
    return packed, cb, t, mpb, nm, _nbins(ctx, packed, t)


@pytest.mark.gpu
def test_c2_front_end_windows_gpu(orc, ctx, c2):
    packed, cb, t, mpb, nm, nb = c2
    for y, x in _windows(np.random.default_rng(22), 2048, 2048, 4):
        e = _front_end_window(orc, ctx, packed, cb, t, mpb, nm, nb, y, x)
        print(f"  window ({y}, {x}): mPb_all {e[0]:.1e}, mPb_nmax {e[1]:.1e}, flips {e[2]}", flush=True)


def _eigen_parity(ctx, nm, out_evals, vec, D):
    """FP64 reference on the same mPb_nmax; returns (eigenvalue rel err, angle, gPb err)."""
    ctx.set_eig_convergence(True)
    ctx.set_eig_options(nvec=9, basis=64, keep=32, tol=1e-5, max_matvecs=60000)
    try:
        ev_ref, vec_ref = ctx.eigs_f64(nm)
    finally:
        ctx.set_eig_convergence(False)
        ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=1e-3)
    rel = float(np.max(np.abs(out_evals[1:] - ev_ref[1:]) / ev_ref[1:]))
    ang = _angle(vec[1:], vec_ref[1:], D)
    spb = ctx.spb_from_evecs(vec, out_evals)
    spb_ref = ctx.spb_from_evecs(vec_ref.astype(np.float32), ev_ref)
    return rel, ang, W_SPB * float(np.abs(spb - spb_ref).max()), ev_ref


@pytest.mark.gpu
def test_c2_eigen_parity_gpu(gpb, ctx, c2):
    packed, cb, _, _, nm, _ = c2
    print(f"C2 whole path at residual <= {PARITY_TOL:g} lambda", flush=True)
    ctx.set_eig_convergence(True)
    ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=PARITY_TOL, max_matvecs=80000)
    try:
        out = ctx.compute(packed, cb)
        st = ctx.stats()
        vec, D = ctx.last_eigenvectors(2048, 2048, 9)
    finally:
        ctx.set_eig_convergence(False)
        ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=1e-3)
    assert np.array_equal(out["mpb_nmax"], nm) and st["converged"] == 9
    ev = np.array(st["evals"][:9])
    rel, ang, gerr, ev_ref = _eigen_parity(ctx, nm, ev, vec, D)
    print(f"  {st['matvecs']} operator applications, eigensolve {st['stage_ms'][4]:.0f} ms; eigenvalues "
          f"{ev_ref[1]:.2e} .. {ev_ref[8]:.2e}: rel err {rel:.1e}, subspace {ang:.1e} rad, gPb {gerr:.1e}", flush=True)
    assert rel <= 1e-5 and ang <= 1e-3 and gerr <= 1e-3


# ---- C3 ---------------------------------------------------------------------------


@pytest.mark.gpu
def test_c3_precision_window_gpu(gpb, ctx):
    print("C3: 1024x1024 window of the 4096x4096 three-class image, production vs all-FP64", flush=True)
    packed, cb = _image(gpb, 4096, 1600, 3, classes=3)
    win = np.ascontiguousarray(packed[1536:2560, 1536:2560])
    ctx.set_eig_convergence(True)
    ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=PARITY_TOL, max_matvecs=80000)
    try:
        out = ctx.compute(win, cb)
        st = ctx.stats()
        vec, D = ctx.last_eigenvectors(1024, 1024, 9)
        ctx.set_eig_options(nvec=9, basis=64, keep=32, tol=1e-5, max_matvecs=60000)
        f64 = ctx.compute_f64(win, cb, with_evecs=False)
    finally:
        ctx.set_eig_convergence(False)
        ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=1e-3)
    nm = out["mpb_nmax"]
    flips = (nm > 0) != (f64["mpb_nmax"] > 0)
    err_nm = np.abs(nm - f64["mpb_nmax"])[~flips].max()
    ev = np.array(st["evals"][:9])
    rel, ang, gerr, ev_ref = _eigen_parity(ctx, nm, ev, vec, D)
    print(f"  mPb_nmax {err_nm:.1e} ({int(flips.sum())} flips); {st['matvecs']} applications; eigenvalues "
          f"{ev_ref[1]:.2e} .. {ev_ref[8]:.2e}: rel err {rel:.1e}, subspace {ang:.1e} rad, gPb {gerr:.1e}", flush=True)
    assert err_nm <= 1e-4 and flips.sum() <= 8
    assert rel <= 1e-5 and ang <= 1e-3 and gerr <= 1e-3
    if flips.sum() == 0:
        # end to end: the same stencil up to rounding, so the whole oriented gPb too
        e2e = np.abs(out["gpb_orient"] - f64["gpb_orient"]).max()
        rel2 = np.max(np.abs(ev[1:] - f64["evals"][1:]) / f64["evals"][1:])
        print(f"  end to end: gPb {e2e:.1e}, eigenvalues {rel2:.1e}", flush=True)
        assert e2e <= 1e-3 and rel2 <= 1e-5


# ---- C4 ---------------------------------------------------------------------------


@pytest.mark.gpu
def test_c4_batched_patches_gpu(gpb, orc, ctx):
    S = 5000
    print("C4 5000x5000, 16 patches, batched Lanczos at residual 1e-3", flush=True)
    packed, cb = _image(gpb, S, 2500, 4)
    rects, cores = gpb.cut_patches(S, S, 4, 16, 4)
    ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=1e-3)
    ctx.compute_patches(packed, cb, rects, cores, mode=1)
    rng = np.random.default_rng(44)
    worst = 0.0
    for p in range(len(rects)):
        x0, y0, w, h = (int(v) for v in rects[p])
        st = ctx.patch_stats(p)
        assert st["converged"] == 9 and st["max_residual"] <= 1e-3
        vec, deg, wup, nm = ctx.patch_eigenvectors(p, w, h, 9)
        full = _full_disc(orc, wup.astype(np.float64))
        D64 = full.sum(0)
        assert np.abs(D64 - deg).max() <= 1e-4 * D64.max()
        ev = st["evals"][:9]
        with ThreadPoolExecutor(9) as ex:
            res = list(ex.map(lambda i: orc.gen_residual(full, D64, ev[i], vec[i].astype(np.float64)), range(9)))
        worst = max(worst, max(res))
        # the tolerance leaves room for the FP32 storage of the returned vectors
        assert max(res) <= 1e-3 * 1.01 + 1e-6, (p, res)
        del full
        if p % 5:
            continue
        # the front end of patches 0, 5, 10 and 15 (a corner, two interior
        # patches, the opposite corner) against the oracle on one window each
        py, px = _windows(rng, h, w, 1)[0]
        sub = np.ascontiguousarray(packed[y0:y0 + h, x0:x0 + w])
        t = ctx.texton(sub, cb)
        _, mpb = ctx.local_cues(sub, t, with_cues=False)
        assert np.array_equal(ctx.nonmax(mpb), nm)
        _front_end_window(orc, ctx, sub, cb, t, mpb, nm, _nbins(ctx, sub, t), py, px)
    print(f"  every patch: 9 pairs, FP64 residual at most {worst:.2e}", flush=True)


# ---- C5 ---------------------------------------------------------------------------


@pytest.mark.gpu
def test_c5_warm_start_gain_gpu(gpb, ctx):
    S = 4096
    print("C5: unseen 4096x4096 evaluation image (seed 5000), plain vs learned vs random block start", flush=True)
    packed, cb = _image(gpb, S, 1600, 5000, classes=3)
    ctx.set_warm_params(np.fromfile(ROOT / "ucm-mpi-gpu_amd" / "weights" / "warm_start.f32", dtype="<f4"))
    ctx.set_eig_options(nvec=9, basis=48, keep=16, tol=1e-3)
    apps = {}
    try:
        for name, mode in (("plain", 0), ("learned", 1), ("random", 2)):
            ctx.set_warm_start(mode, 80, 24)
            ctx.compute(packed, cb)
            st = ctx.stats()
            apps[name] = st["matvecs"]
            print(f"  {name}: {st['matvecs']} operator applications, eigensolve {st['stage_ms'][4]:.0f} ms, fallback "
                  f"{st['fallback']}, residual {st['max_residual']:.2e}", flush=True)
            assert st["max_residual"] <= 1e-3 and st["fallback"] == 0 and st["warm_mode"] == mode
    finally:
        ctx.set_warm_start(0)
    assert apps["learned"] < apps["plain"] and apps["learned"] < apps["random"]

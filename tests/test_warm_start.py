"""Learned warm start (no reference counterpart): network, loss, gradients,
cue features, the warm-started eigensolve and its fallback.

CPU: the FP64 restatement (oracle/warm_oracle.py) is pinned by central finite
differences of its loss, by scipy's principal angles, by the loss's
invariances (rotation, scale and constant shift of the predicted block) and
its zero at the target subspace, and by BF16 rounding cases.  GPU: the
MFMA forward / backward kernels against the restatement with BF16 rounding at
the same points (loss to 1e-4 relative, output to 1e-2 of its scale, every
layer's weight gradient to 2e-2 relative L2), the cue features against the
oracle's cue maps, the warm-started solve against plain Lanczos at a tight
residual, the fallback when the network's block is degenerate, and training
lowering the loss on held-out images.
"""
import importlib.util
import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sl

from conftest import ROOT, synth_packed


def _warm_oracle():
    name = "gpb_warm_oracle"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, ROOT / "oracle" / "warm_oracle.py")
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    return sys.modules[name]


def _subspace_case(rng, n, h, w, noise):
    """Targets from D-orthonormal V (D-orthogonal to the constant) and a
    prediction P = V R + noise."""
    Db = rng.uniform(0.5, 2.0, (n, h, w))
    V = rng.normal(size=(n, h, w, 8))
    for i in range(n):
        d = Db[i].reshape(-1)
        v = V[i].reshape(-1, 8)
        v -= (d[:, None] * v).sum(0) / d.sum()
        L = np.linalg.cholesky(v.T @ (d[:, None] * v))
        V[i] = (v @ np.linalg.inv(L).T).reshape(h, w, 8)
    T = Db[..., None] * V
    P = V @ rng.normal(size=(8, 8)) + noise * rng.normal(size=V.shape)
    return P, T, Db, V


def test_bf16_rounding_cases():
    wo = _warm_oracle()
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40])
    y = wo.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0  # tie to even
    assert y[2] == 1 + 2 ** -6  # tie to even (upwards)
    assert y[3] == 1 + 2 ** -7 and y[4] == -2.5
    assert np.isfinite(y[5]) and abs(y[6]) < 1e-39
    # every BF16 value is a fixed point
    r = wo.bf16(np.random.default_rng(0).normal(size=1000))
    assert np.array_equal(wo.bf16(r), r)


def test_oracle_loss_is_principal_angles():
    wo = _warm_oracle()
    rng = np.random.default_rng(5)
    P, T, Db, V = _subspace_case(rng, 2, 16, 16, noise=0.7)
    loss, _ = wo.loss_and_grad(P, T, Db)
    ref = 0.0
    for i in range(2):
        s = np.sqrt(Db[i].reshape(-1))[:, None]
        p = P[i].reshape(-1, 8)
        d = Db[i].reshape(-1)
        p = p - (d[:, None] * p).sum(0) / d.sum()
        ang = sl.subspace_angles(s * p, s * V[i].reshape(-1, 8))
        ref += np.mean(np.sin(ang) ** 2)
    assert abs(loss - ref / 2) < 1e-8
    # zero at the exact subspace; invariant to rotation, scale and a constant shift
    l0, _ = wo.loss_and_grad(V, T, Db)
    assert l0 < 1e-8  # the 1e-9 tr(G) / 8 ridge
    R = rng.normal(size=(8, 8))
    l1, _ = wo.loss_and_grad(P @ R * 3.0 + 5.0, T, Db)
    assert abs(l1 - loss) < 1e-7  # up to the trace-scaled ridge


def test_oracle_loss_gradient_vs_finite_differences():
    wo = _warm_oracle()
    rng = np.random.default_rng(7)
    P, T, Db, _ = _subspace_case(rng, 2, 8, 8, noise=1.0)
    _, dP = wo.loss_and_grad(P, T, Db)
    for _ in range(12):
        idx = tuple(rng.integers(0, s) for s in P.shape)
        e = np.zeros_like(P)
        e[idx] = 1e-6
        fd = (wo.loss_and_grad(P + e, T, Db)[0] - wo.loss_and_grad(P - e, T, Db)[0]) / 2e-6
        assert abs(fd - dP[idx]) <= 1e-6 * max(1.0, abs(dP).max()), (idx, fd, dP[idx])


def test_oracle_network_gradient_vs_finite_differences():
    wo = _warm_oracle()
    rng = np.random.default_rng(11)
    params = wo.init_params(3)
    offs, _ = wo.offsets()
    # nonzero biases so that every bias gradient path is exercised
    for l, ((cin, cout, _), (w_off, b_off)) in enumerate(zip(wo.LAYERS, offs)):
        live = wo.N_OUT if l == len(wo.LAYERS) - 1 else cout
        params[b_off:b_off + live] = rng.normal(0, 0.05, live)
    x = rng.normal(size=(2, 8, 8, wo.N_IN))
    P, T, Db, _ = _subspace_case(rng, 2, 8, 8, noise=1.0)
    loss, grad, _ = wo.loss_grad(params, x, T, Db)

    def L(p):
        return wo.loss_grad(p, x, T, Db)[0]

    eps = 1e-6
    for l, ((cin, cout, _), (w_off, b_off)) in enumerate(zip(wo.LAYERS, offs)):
        live = min(cout, 8) if l == len(wo.LAYERS) - 1 else cout
        picks = [w_off + int(rng.integers(0, 9)) * cout * cin + int(rng.integers(0, live)) * cin +
                 int(rng.integers(0, cin)) for _ in range(2)] + [b_off + int(rng.integers(0, live))]
        scale = max(np.abs(grad[w_off:b_off + cout]).max(), 1e-12)
        for i in picks:
            e = np.zeros_like(params)
            e[i] = eps
            fd = (L(params + e) - L(params - e)) / (2 * eps)
            assert abs(fd - grad[i]) <= 1e-4 * scale, (l, i, fd, grad[i], scale)


@pytest.mark.gpu
def test_warm_loss_grad_gpu_vs_oracle(gpb, ctx):
    wo = _warm_oracle()
    print("warm start: GPU loss / gradient vs the restatement (2 x 16 x 24 cells)", flush=True)
    rng = np.random.default_rng(21)
    ctx.warm_init(5)
    params = ctx.warm_params().astype(np.float64)
    # random biases (the GPU keeps the head's channels 8..31 at zero)
    offs, _ = wo.offsets()
    for l, ((cin, cout, _), (w_off, b_off)) in enumerate(zip(wo.LAYERS, offs)):
        live = min(cout, 8) if l == len(wo.LAYERS) - 1 else cout
        params[b_off:b_off + live] = rng.normal(0, 0.05, live)
    ctx.set_warm_params(params.astype(np.float32))
    params = ctx.warm_params().astype(np.float64)
    n, h, w = 2, 16, 24
    x = rng.uniform(0, 1, size=(n, h, w, wo.N_IN))
    _, T, Db, _ = _subspace_case(rng, n, h, w, noise=0.0)
    loss_g, grad_g, y_g = ctx.warm_loss_grad(x.astype(np.float32), T.astype(np.float32), Db.astype(np.float32))
    loss_o, grad_o, y_o = wo.loss_grad(params, x.astype(np.float32).astype(np.float64),
                                       T.astype(np.float32).astype(np.float64),
                                       Db.astype(np.float32).astype(np.float64), rounding="bf16")
    print(f"  loss gpu {loss_g:.6f} oracle {loss_o:.6f}", flush=True)
    assert abs(loss_g - loss_o) <= 1e-4 * max(1.0, abs(loss_o))
    assert np.abs(y_g - y_o).max() <= 1e-2 * np.abs(y_o).max()
    for l, ((cin, cout, _), (w_off, b_off)) in enumerate(zip(wo.LAYERS, offs)):
        g0, g1 = grad_o[w_off:b_off + cout], grad_g[w_off:b_off + cout]
        rel = np.linalg.norm(g1 - g0) / max(np.linalg.norm(g0), 1e-30)
        print(f"  layer {l}: grad rel err {rel:.2e} (|g| {np.linalg.norm(g0):.3e})", flush=True)
        assert rel <= 2e-2, l


@pytest.mark.gpu
def test_warm_features_gpu_vs_oracle(gpb, orc, ctx):
    W, H = 70, 53
    print(f"warm start: cue features {W}x{H}", flush=True)
    rgb, packed = synth_packed(gpb, W, H, regions=8, seed=13)
    cb = gpb.make_codebook(packed, seed=13)
    f = ctx.warm_features(packed, cb)
    t = orc.texton(packed, cb)
    cues, _ = orc.local_cues(packed, t)
    wts = np.array([0.0146, 0.0145, 0.0163, 0.0210, 0.0243, 0.0287, 0.0166, 0.0185, 0.0204, 0.0101, 0.0111, 0.0141])
    hq, wq = (H + 3) // 4, (W + 3) // 4
    ref = np.zeros((32, hq, wq))
    for ch in range(4):
        contrib = sum(wts[3 * ch + s] * cues[3 * ch + s] for s in range(3))  # [8, H, W]
        for cy in range(hq):
            for cx in range(wq):
                ref[8 * ch:8 * ch + 8, cy, cx] = contrib[:, 4 * cy:4 * cy + 4, 4 * cx:4 * cx + 4].mean(axis=(1, 2))
    err = np.abs(f - ref).max()
    print(f"  max abs err {err:.2e} (feature max {ref.max():.3f})", flush=True)
    assert err <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2])
def test_warm_solve_matches_plain_gpu(gpb, ctx, mode):
    W, H = 192, 160
    print(f"warm start: mode {mode} solve vs plain {W}x{H}", flush=True)
    rgb, packed = synth_packed(gpb, W, H, regions=14, seed=17)
    cb = gpb.make_codebook(packed, seed=17)
    ctx.set_eig_options(nvec=9, basis=40, keep=20, tol=1e-6)
    ref = ctx.compute(packed, cb)
    ev0 = np.array(ctx.stats()["evals"][:9])
    try:
        ctx.warm_init(9)
        ctx.set_warm_start(mode, basis=64, keep=24)
        out = ctx.compute(packed, cb)
        st = ctx.stats()
    finally:
        ctx.set_warm_start(0)
    ev = np.array(st["evals"][:9])
    print(f"  matvecs {st['matvecs']} block {st['block']} fallback {st['fallback']} residual "
          f"{st['max_residual']:.2e} warm ms {st['warm_ms']:.2f}", flush=True)
    assert st["warm_mode"] == mode and st["max_residual"] <= 1e-6
    rel = np.abs(ev[1:] - ev0[1:]) / ev0[1:]
    print(f"  eigenvalue rel diff {rel.max():.2e}", flush=True)
    assert rel.max() <= 5e-5
    assert np.abs(out["gpb_max"] - ref["gpb_max"]).max() <= 1e-3


@pytest.mark.gpu
def test_warm_degenerate_block_falls_back_gpu(gpb, ctx):
    """A network with all-zero parameters predicts a rank-deficient block:
    the block solve breaks down and plain Lanczos answers."""
    W, H = 128, 96
    print(f"warm start: forced fallback {W}x{H}", flush=True)
    rgb, packed = synth_packed(gpb, W, H, regions=10, seed=19)
    cb = gpb.make_codebook(packed, seed=19)
    ctx.set_eig_options(nvec=9, basis=40, keep=20, tol=1e-4)
    ref = ctx.compute(packed, cb)
    ev0 = np.array(ctx.stats()["evals"][:9])
    try:
        ctx.set_warm_params(np.zeros(gpb.warm_param_count(), np.float32))
        ctx.set_warm_start(1)
        ctx.compute(packed, cb)
        st = ctx.stats()
    finally:
        ctx.set_warm_start(0)
    print(f"  fallback {st['fallback']} matvecs {st['matvecs']} residual {st['max_residual']:.2e}", flush=True)
    assert st["fallback"] == 1 and st["max_residual"] <= 1e-4
    ev = np.array(st["evals"][:9])
    assert np.max(np.abs(ev[1:] - ev0[1:]) / ev0[1:]) <= 1e-3


@pytest.mark.gpu
def test_warm_training_lowers_held_out_loss_gpu(gpb, ctx):
    W = H = 96
    n_train, n_test = 24, 6
    print(f"warm start: training on {n_train} images {W}x{H}", flush=True)
    ctx.set_eig_options(nvec=9, basis=40, keep=20, tol=1e-4)
    ctx.warm_init(1)
    ctx.warm_dataset_reset(W, H, n_train + n_test)
    for i in range(n_train + n_test):
        rgb, packed = synth_packed(gpb, W, H, regions=8 + i % 8, seed=100 + i)
        ctx.warm_dataset_add(packed, gpb.make_codebook(packed, seed=100 + i))
    assert ctx.warm_dataset_size() == n_train + n_test
    before = ctx.warm_eval_loss(n_train, n_test)
    losses = ctx.warm_train(0, n_train, 40, 8, 2e-3, 3, augment=True)
    after = ctx.warm_eval_loss(n_train, n_test)
    print(f"  held-out loss {before:.4f} -> {after:.4f}; training loss {losses[0]:.4f} -> {losses[-1]:.4f}",
          flush=True)
    assert losses[-1] < losses[0]
    assert after < before


@pytest.mark.gpu
@pytest.mark.parametrize("tf", [1, 2, 4, 7])
def test_feature_symmetries_gpu(gpb, ctx, tf):
    """The symmetry augmentation's orientation map, checked physically: for
    every orientation o of a flipped / transposed image, the best-matching
    orientation of the original's features (moved with the geometry) is
    pi(o), with o sent to -o by a flip and to pi/2 - o by a transpose.  The
    match is not exact: the half-discs' half-pixel origin makes the cue maps
    only approximately mirror-symmetric (5-13 % mean difference even for the
    orientations that map to themselves)."""
    S = 96
    print(f"warm start: feature symmetry {tf} on {S}x{S}", flush=True)
    rgb, packed = synth_packed(gpb, S, S, regions=9, seed=23)
    cb = gpb.make_codebook(packed, seed=23)
    f0 = ctx.warm_features(packed, cb)  # [32, 24, 24]
    img = rgb
    if tf & 1:
        img = img[:, ::-1]
    if tf & 2:
        img = img[::-1, :]
    if tf & 4:
        img = img.transpose(1, 0, 2)
    f1 = ctx.warm_features(gpb.pack_rgb(np.ascontiguousarray(img)), cb)

    def moved(m):  # the original's map seen from the transformed image: transpose . flip_y . flip_x
        if tf & 4:
            m = m.T
        if tf & 2:
            m = m[::-1, :]
        if tf & 1:
            m = m[:, ::-1]
        return m

    worst = 0.0
    for o in range(8):
        so = o
        if tf & 1:
            so = (8 - so) & 7
        if tf & 2:
            so = (8 - so) & 7
        if tf & 4:
            so = (12 - so) & 7
        errs = []
        for cand in range(8):
            e = sum(np.abs(f1[8 * c + o] - moved(f0[8 * c + cand]))[2:-2, 2:-2].mean() for c in range(4))
            errs.append(e)
        ref = sum(np.abs(moved(f0[8 * c + so]))[2:-2, 2:-2].mean() for c in range(4))
        worst = max(worst, errs[so] / ref)
        print(f"  o {o}: expected source {so}, best {int(np.argmin(errs))}, rel diff {errs[so] / ref:.3f}", flush=True)
        assert int(np.argmin(errs)) == so, o
    assert worst < 0.25

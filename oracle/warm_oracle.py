"""CPU restatement (numpy, FP64) of the learned warm start's network, loss and
gradients.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels (ucm-mpi-gpu_amd/csrc/k_warm.hip).  The product never imports it.

The warm start has no counterpart in the reference (its eigensolvers start
from a fixed or random vector: spectralPb.cpp:221-244,
damascene/src/noReorthog/lanczos.cu:393-752), so nothing here restates
reference code.  What pins it: the gradient is checked against central finite
differences of the loss (tests/test_warm_start.py), and the loss against
scipy's principal angles between subspaces, its invariances (sign, scale,
rotation of the predicted block) and its zero at the exact subspace.

Network (cells h x w, h and w multiples of 8; channels in brackets):
  x0[64] -> a0 = relu(c0)[32] -> s0 = relu(c1)[32]
  p1 = pool(s0) -> a1 = relu(c2)[64] -> s1 = relu(c3)[64]
  p2 = pool(s1) -> a2 = relu(c4)[128] -> s2 = relu(c5)[128]
  p3 = pool(s2) -> a3 = relu(c6)[128] -> b3 = relu(c7)[128]
  u2 = up(b3) + s2 -> d2 = relu(c8)[64]
  u1 = up(d2) + s1 -> d1 = relu(c9)[32]
  u0 = up(d1) + s0 -> d0 = relu(c10)[32] -> y = c11[32], block = y[..., :8]
Conv weights W[9][cout][cin] (tap t = 3 (dy + 1) + (dx + 1)): out[y, x] =
b + sum_t in[y + dy, x + dx] W[t]^T with zero padding; pool = 2x2 mean,
up = 2x nearest.  Parameter vector: per layer W then b.

With rounding="bf16" the activations, the input, the weights and the
backward gradients are rounded to BF16 (round to nearest even) at the points
where the GPU stores them; otherwise all arithmetic is FP64.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(64, 32, 1), (32, 32, 1), (32, 64, 1), (64, 64, 1), (64, 128, 1), (128, 128, 1), (128, 128, 1),
          (128, 128, 1), (128, 64, 1), (64, 32, 1), (32, 32, 1), (32, 32, 0)]
N_IN = 34  # 32 cue features + 2 coordinates (the network input pads to 64)
N_OUT = 8


def offsets():
    """[(w_off, b_off)] per layer and the total parameter count."""
    out, at = [], 0
    for cin, cout, _ in LAYERS:
        out.append((at, at + 9 * cin * cout))
        at += 9 * cin * cout + cout
    return out, at


def unpack(params: np.ndarray):
    offs, total = offsets()
    assert params.size == total
    Ws, bs = [], []
    for (cin, cout, _), (wo, bo) in zip(LAYERS, offs):
        Ws.append(params[wo:wo + 9 * cin * cout].reshape(9, cout, cin).astype(np.float64))
        bs.append(params[bo:bo + cout].astype(np.float64))
    return Ws, bs


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (nearest even) and back to FP64 (via FP32)."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(x):
    return x


def _taps():
    return [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv(x, W, b=None):
    n, h, w, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((n, h, w, W.shape[1]))
    for t, (dy, dx) in enumerate(_taps()):
        out += xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :] @ W[t].T
    if b is not None:
        out += b
    return out


def conv_dgrad(g, W):
    n, h, w, _ = g.shape
    gp = np.pad(g, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((n, h, w, W.shape[2]))
    for t, (dy, dx) in enumerate(_taps()):
        out += gp[:, 1 - dy:1 - dy + h, 1 - dx:1 - dx + w, :] @ W[t]
    return out


def conv_wgrad(x, g):
    n, h, w, cin = x.shape
    cout = g.shape[3]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    gm = g.reshape(-1, cout)
    dW = np.zeros((9, cout, cin))
    for t, (dy, dx) in enumerate(_taps()):
        dW[t] = gm.T @ xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :].reshape(-1, cin)
    return dW, gm.sum(0)


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def upsum(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def forward(params: np.ndarray, x: np.ndarray, rounding: str | None = None):
    """x [n, h, w, 34] -> (cache, y [n, h, w, 32])."""
    r = bf16 if rounding == "bf16" else _ident
    Ws, bs = unpack(params)
    if rounding == "bf16":
        Ws = [bf16(W) for W in Ws]
    n, h, w, _ = x.shape
    x0 = np.zeros((n, h, w, 64))
    x0[..., :N_IN] = r(x.astype(np.float64))

    def cr(inp, l):
        o = conv(inp, Ws[l], bs[l])
        return r(np.maximum(o, 0.0)) if LAYERS[l][2] else o

    c = dict(x0=x0)
    c["a0"] = cr(x0, 0)
    c["s0"] = cr(c["a0"], 1)
    c["p1"] = r(pool(c["s0"]))
    c["a1"] = cr(c["p1"], 2)
    c["s1"] = cr(c["a1"], 3)
    c["p2"] = r(pool(c["s1"]))
    c["a2"] = cr(c["p2"], 4)
    c["s2"] = cr(c["a2"], 5)
    c["p3"] = r(pool(c["s2"]))
    c["a3"] = cr(c["p3"], 6)
    c["b3"] = cr(c["a3"], 7)
    c["u2"] = r(up(c["b3"]) + c["s2"])
    c["d2"] = cr(c["u2"], 8)
    c["u1"] = r(up(c["d2"]) + c["s1"])
    c["d1"] = cr(c["u1"], 9)
    c["u0"] = r(up(c["d1"]) + c["s0"])
    c["d0"] = cr(c["u0"], 10)
    y = cr(c["d0"], 11)
    c["Ws"] = Ws
    return c, y


def loss_and_grad(P: np.ndarray, T: np.ndarray, Db: np.ndarray):
    """Mean over images of 1 - tr(G^-1 C C^T) / 8 and its gradient w.r.t. P.

    P, T [n, h, w, 8], Db [n, h, w].  Per image, with cells as rows:
    P' = P - 1 m^T (m = Db-weighted mean), G = P'^T diag(Db) P' (plus a
    1e-9 tr(G) / 8 ridge), C = P'^T T, the loss is the mean squared sine of the
    principal angles between span(P') and span(T-implied subspace).
    """
    n = P.shape[0]
    loss = 0.0
    dP = np.zeros_like(P, dtype=np.float64)
    for i in range(n):
        p = P[i].reshape(-1, 8).astype(np.float64)
        t = T[i].reshape(-1, 8).astype(np.float64)
        d = Db[i].reshape(-1).astype(np.float64)
        S = d.sum()
        m = (d[:, None] * p).sum(0) / S
        pc = p - m
        G = pc.T @ (d[:, None] * pc)
        G += np.eye(8) * (1e-9 * np.trace(G) / 8)
        Cm = pc.T @ t
        A = Cm @ Cm.T
        Gi = np.linalg.inv(G)
        f = np.trace(Gi @ A)
        loss += 1.0 - f / 8.0
        B = Gi @ A @ Gi
        dpc = -0.25 * (t @ Cm.T @ Gi - d[:, None] * (pc @ B))
        dp = dpc - d[:, None] * dpc.sum(0) / S
        dP[i] = dp.reshape(P[i].shape)
    return loss / n, dP / n


def backward(params: np.ndarray, cache: dict, dy: np.ndarray, rounding: str | None = None) -> np.ndarray:
    """Gradient of the loss w.r.t. the parameter vector, given dL/dy (32 channels)."""
    r = bf16 if rounding == "bf16" else _ident
    Ws = cache["Ws"]
    offs, total = offsets()
    grad = np.zeros(total)
    c = cache

    def put(l, x, g):
        dW, db = conv_wgrad(x, g)
        wo, bo = offs[l]
        grad[wo:wo + dW.size] = dW.reshape(-1)
        grad[bo:bo + db.size] = db

    def mask(g, a):
        return g * (a > 0)

    dy = r(dy)
    put(11, c["d0"], dy)
    g_d0 = r(mask(conv_dgrad(dy, Ws[11]), c["d0"]))
    put(10, c["u0"], g_d0)
    g_u0 = r(conv_dgrad(g_d0, Ws[10]))
    g_d1 = r(mask(upsum(g_u0), c["d1"]))
    put(9, c["u1"], g_d1)
    g_u1 = r(conv_dgrad(g_d1, Ws[9]))
    g_d2 = r(mask(upsum(g_u1), c["d2"]))
    put(8, c["u2"], g_d2)
    g_u2 = r(conv_dgrad(g_d2, Ws[8]))
    g_b3 = r(mask(upsum(g_u2), c["b3"]))
    put(7, c["a3"], g_b3)
    g_a3 = r(mask(conv_dgrad(g_b3, Ws[7]), c["a3"]))
    put(6, c["p3"], g_a3)
    g_p3 = r(conv_dgrad(g_a3, Ws[6]))
    g_s2 = r(mask(g_u2 + 0.25 * up(g_p3), c["s2"]))
    put(5, c["a2"], g_s2)
    g_a2 = r(mask(conv_dgrad(g_s2, Ws[5]), c["a2"]))
    put(4, c["p2"], g_a2)
    g_p2 = r(conv_dgrad(g_a2, Ws[4]))
    g_s1 = r(mask(g_u1 + 0.25 * up(g_p2), c["s1"]))
    put(3, c["a1"], g_s1)
    g_a1 = r(mask(conv_dgrad(g_s1, Ws[3]), c["a1"]))
    put(2, c["p1"], g_a1)
    g_p1 = r(conv_dgrad(g_a1, Ws[2]))
    g_s0 = r(mask(g_u0 + 0.25 * up(g_p1), c["s0"]))
    put(1, c["a0"], g_s0)
    g_a0 = r(mask(conv_dgrad(g_s0, Ws[1]), c["a0"]))
    put(0, c["x0"], g_a0)
    return grad


def loss_grad(params: np.ndarray, x: np.ndarray, T: np.ndarray, Db: np.ndarray, rounding: str | None = None):
    """(loss, parameter gradient, block y[..., :8]) for one batch."""
    cache, y = forward(params, x, rounding)
    loss, dP = loss_and_grad(y[..., :N_OUT], T, Db)
    dy = np.zeros_like(y)
    dy[..., :N_OUT] = dP
    if rounding == "bf16":
        dy = dy.astype(np.float32).astype(np.float64)
    return loss, backward(params, cache, dy, rounding), y[..., :N_OUT]


def init_params(seed: int) -> np.ndarray:
    """He-scaled normal weights (head: half the variance), zero biases, and the
    head's output channels 8..31 zero (the layout the GPU keeps)."""
    rng = np.random.default_rng(seed)
    offs, total = offsets()
    p = np.zeros(total)
    for l, ((cin, cout, _), (wo, bo)) in enumerate(zip(LAYERS, offs)):
        sd = np.sqrt((1.0 if l == len(LAYERS) - 1 else 2.0) / (9 * cin))
        p[wo:wo + 9 * cin * cout] = rng.normal(0.0, sd, 9 * cin * cout)
    cin, cout, _ = LAYERS[-1]
    wo, bo = offs[-1]
    W = p[wo:wo + 9 * cin * cout].reshape(9, cout, cin)
    W[:, N_OUT:, :] = 0.0
    p[bo + N_OUT:bo + cout] = 0.0
    return p

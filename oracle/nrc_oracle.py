"""Radiance-cache oracle: FP64 numpy restatement of the hash-grid encoding,
the 32 -> 5 x 64 ReLU -> 3 MLP, the relative L2 loss and their gradients.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
bench's CPU baseline as the checker; the product never calls it.

The engine has no radiance cache, so nothing in the reference pins these
functions; they are pinned by closed forms and by finite differences in
tests/test_nrc.py:
  - level_resolutions: closed form (base * b^l, b^(levels-1) = max / base);
  - encode: trilinear interpolation reproduces a linear function of the
    lattice exactly; a query on a lattice vertex returns that vertex's
    (hashed or dense) entry;
  - forward: a hand-built network with one path of non-zero weights gives
    the product of those weights;
  - loss_and_grads: central finite differences of the loss with the
    denominator frozen (the stop-gradient), for weights and table entries.

Two numeric modes.  quantize=False is plain FP64 over the given parameters.
quantize=True rounds every MFMA operand to bfloat16 as the kernels do (the
inputs of each layer, the back-propagated dZ of each layer and the weights)
while accumulating in FP64, so that GPU and oracle differ only by FP32
accumulation order.
"""
import math

import numpy as np

K = [32, 64, 64, 64, 64, 64]
N = [64, 64, 64, 64, 64, 3]
MASTER_OFFSETS = np.cumsum([0] + [k * n for k, n in zip(K, N)])
MASTER_SIZE = int(MASTER_OFFSETS[-1])
LOSS_FLOOR = 0.01
BETA1, BETA2, EPS = 0.9, 0.99, 1e-15
PRIME_Y, PRIME_Z = np.uint32(2654435761), np.uint32(805459861)


def bf16_round(x):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def level_resolutions(levels, base, maxres):
    """Cells per axis of each level: floor(base * b^l), b = (max / base)^(1 / (levels - 1))."""
    b = math.exp((math.log(maxres) - math.log(base)) / (levels - 1)) if levels > 1 else 1.0
    return np.array([math.floor(base * b ** l + 1e-9) for l in range(levels)], dtype=np.float32)


class Params:
    """Cache configuration (atlas_rt_nrc_params)."""

    def __init__(self, bounds_min, bounds_max, levels=12, log2_table=19, base=16, maxres=2048):
        self.bmin = np.asarray(bounds_min, np.float32)
        self.binv = (1.0 / (np.asarray(bounds_max, np.float64) - np.asarray(bounds_min, np.float64))).astype(np.float32)
        self.levels, self.log2_table = levels, log2_table
        self.T = 1 << log2_table
        self.res = level_resolutions(levels, base, maxres)
        self.dense = [(float(r) + 1) ** 3 <= self.T for r in self.res]


def corners(P, queries, lev):
    """(idx (n, 8) uint32 entry within the level, w (n, 8) float64) for level `lev`.
    Index arithmetic in float32 / uint32 as on the GPU."""
    pos = queries[:, 0:3].astype(np.float32)
    x = np.clip((pos - P.bmin) * P.binv, np.float32(0), np.float32(1)).astype(np.float32)
    x = np.where(np.isnan(x), np.float32(0), x)
    R = np.float32(P.res[lev])
    s = (x * R).astype(np.float32)
    c = np.minimum(np.floor(s), R - np.float32(1)).astype(np.float32)
    f = (s - c).astype(np.float32)
    ic = c.astype(np.uint32)
    n = len(queries)
    idx = np.zeros((n, 8), np.uint32)
    w = np.zeros((n, 8), np.float64)
    R1 = np.uint32(int(R) + 1)
    for k in range(8):
        o = np.array([k & 1, (k >> 1) & 1, k >> 2], np.uint32)
        p = ic + o
        if P.dense[lev]:
            h = p[:, 0] + R1 * (p[:, 1] + R1 * p[:, 2])
        else:
            h = (p[:, 0] ^ (p[:, 1] * PRIME_Y) ^ (p[:, 2] * PRIME_Z)) & np.uint32(P.T - 1)
        idx[:, k] = h
        wk = np.ones(n, np.float64)
        for a in range(3):
            wk = wk * np.where(o[a] == 1, f[:, a].astype(np.float64), 1.0 - f[:, a].astype(np.float64))
        w[:, k] = wk
    return idx, w


def encode(P, table, queries):
    """32 inputs per query.  table: (levels, T, 2) float64 (FP16 values).
    Returns (X (n, 32), per-level corners)."""
    n = len(queries)
    X = np.zeros((n, 32), np.float64)
    cs = []
    for lev in range(P.levels):
        idx, w = corners(P, queries, lev)
        cs.append((idx, w))
        X[:, 2 * lev] = (w * table[lev, idx, 0]).sum(axis=1)
        X[:, 2 * lev + 1] = (w * table[lev, idx, 1]).sum(axis=1)
    q = queries.astype(np.float64)
    X[:, 24:27] = q[:, 4:7]
    X[:, 27:30] = q[:, 8:11]
    X[:, 30] = q[:, 3]
    X[:, 31] = 1.0
    return X, cs


def split_weights(master):
    return [np.asarray(master[MASTER_OFFSETS[l]:MASTER_OFFSETS[l + 1]], np.float64).reshape(K[l], N[l])
            for l in range(6)]


def forward(W, X, quantize=False):
    """Layer inputs [X0..X5] and outputs y (n, 3)."""
    q = bf16_round if quantize else (lambda a: a)
    Wq = [q(w) for w in W]
    xs = []
    h = X
    for l in range(6):
        h = q(h)
        xs.append(h)
        z = h @ Wq[l]
        h = np.maximum(z, 0.0) if l < 5 else z
    return xs, h, Wq


def loss(y, t):
    den = y * y + LOSS_FLOOR
    return float(((y - t) ** 2 / den).sum() / (3 * len(y)))


def loss_and_grads(P, master, table, queries, targets, quantize=False, frozen_den=None):
    """Relative L2 loss and its gradients w.r.t. the master weights (MASTER_SIZE,)
    and the table (levels, T, 2).  The denominator is a stop-gradient; with
    frozen_den given it is used in place of y^2 + 0.01 (finite differences)."""
    W = split_weights(master)
    X, cs = encode(P, table, queries)
    xs, y, Wq = forward(W, X, quantize)
    t = targets[:, :3].astype(np.float64)
    den = y * y + LOSS_FLOOR if frozen_den is None else frozen_den
    n = len(y)
    L = float(((y - t) ** 2 / den).sum() / (3 * n))
    q = bf16_round if quantize else (lambda a: a)
    dz = q(2.0 * (y - t) / den / (3 * n))
    grads = np.zeros(MASTER_SIZE, np.float64)
    for l in range(5, -1, -1):
        grads[MASTER_OFFSETS[l]:MASTER_OFFSETS[l + 1]] = (xs[l].T @ dz).reshape(-1)
        dx = dz @ Wq[l].T
        if l > 0:
            dz = q(dx * (xs[l] > 0))
    gt = np.zeros((P.levels, P.T, 2), np.float64)
    for lev, (idx, w) in enumerate(cs):
        for c in range(2):
            np.add.at(gt[lev, :, c], idx.reshape(-1), (w * dx[:, 2 * lev + c:2 * lev + c + 1]).reshape(-1))
    return L, grads, gt, y


def predict(P, master, table, queries, quantize=False):
    X, _ = encode(P, table, queries)
    return forward(split_weights(master), X, quantize)[1]


def adam_first_step(p, g, lr):
    """Adam's first step from zero moments: lr * sqrt(1 - b2) / (1 - b1) = lr, and
    m / (sqrt(v) + eps) = sign(g) up to eps, so p - lr * sign(g) (g = 0: unchanged)."""
    g = np.asarray(g, np.float64)
    m = (1 - BETA1) * g
    v = (1 - BETA2) * g * g
    return p - lr * m / (np.sqrt(v) + EPS)


def half_table(bits, levels, log2_table):
    """FP16 table bits (uint16/int16) -> (levels, T, 2) float64."""
    return np.asarray(bits).view(np.float16).astype(np.float64).reshape(levels, 1 << log2_table, 2)

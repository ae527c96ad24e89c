"""CPU restatement of the learned radiance cache (test infrastructure only).

The cache has no counterpart in the reference: it follows published methods,
restated here so that the HIP kernels can be checked against it.
- Multiresolution hash encoding: Mueller, Evans, Schied, Keller, "Instant
  Neural Graphics Primitives with a Multiresolution Hash Encoding" (2022),
  sec. 3: L levels of resolution N_l = floor(N_min * b^l), b = exp((ln N_max -
  ln N_min) / (L - 1)); per level the 8 corners of the cell around x*N_l + 0.5,
  trilinear weights, a dense index where the (N_l + 2)^3 grid fits the table and
  the spatial hash (x*1 xor y*2654435761 xor z*805459861) mod T otherwise.
- MLP: 3 hidden layers of width 64 with ReLU plus a linear output layer of 3,
  no biases (the "4-layer" network: four weight matrices). Inputs: the 32
  encoded features, the voxel normal (3) and albedo (3), zero-padded to 48.
- Loss L = (1/B) sum_s 1/2 |y_s - t_s|^2.
- Adam: Kingma & Ba, "Adam: A Method for Stochastic Optimization" (2015),
  algorithm 1, with bias correction.

`bf16=True` rounds the operands that the GPU keeps in bfloat16 (table entries,
weights, layer inputs and back-propagated gradients) to nearest-even bfloat16,
and accumulates in float64; `bf16=False` is the exact float64 model used for
the finite-difference check.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

PRIMES = (1, 2654435761, 805459861)
IN_PAD = 48
OUT_PAD = 16


@dataclass
class Config:
    levels: int = 16
    features: int = 2
    log2_table: int = 19
    n_min: int = 16
    n_max: int = 256
    width: int = 64
    extent: float = 128.0        # voxel units mapped to [0, 1)
    lr: float = 1e-2
    beta1: float = 0.9
    beta2: float = 0.99
    eps: float = 1e-10

    @property
    def table(self) -> int:
        return 1 << self.log2_table


def level_resolution(cfg: Config, level: int) -> int:
    if cfg.levels == 1:
        return cfg.n_min
    b = np.exp((np.log(np.float64(cfg.n_max)) - np.log(np.float64(cfg.n_min))) / (cfg.levels - 1))
    return int(np.floor(np.float64(cfg.n_min) * b ** level))


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)) << np.uint64(16)
    out = (r & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(a), np.nan, out)


def corner_index(cfg: Config, level: int, c: np.ndarray) -> np.ndarray:
    """Table index of integer corners c (..., 3)."""
    N = level_resolution(cfg, level)
    T = cfg.table
    c = c.astype(np.uint64)
    if (N + 2) ** 3 <= T:
        s = np.uint64(N + 2)
        idx = c[..., 0] + c[..., 1] * s + c[..., 2] * s * s
    else:
        m = np.uint64(0xFFFFFFFF)
        idx = ((c[..., 0] * np.uint64(PRIMES[0])) & m) ^ ((c[..., 1] * np.uint64(PRIMES[1])) & m) ^ \
              ((c[..., 2] * np.uint64(PRIMES[2])) & m)
    return (idx % np.uint64(T)).astype(np.int64)


def encode(cfg: Config, table: np.ndarray, pos: np.ndarray, bf16_table: bool = False):
    """table (L, T, F); pos (B, 3) voxel units. Returns (features (B, L*F),
    corner indices (L, 8, B), corner weights (L, 8, B))."""
    B = pos.shape[0]
    x = np.asarray(pos, np.float64) / np.float64(cfg.extent)
    tab = bf16(table) if bf16_table else np.asarray(table, np.float64)
    feats = np.zeros((B, cfg.levels * cfg.features))
    idxs = np.zeros((cfg.levels, 8, B), np.int64)
    wts = np.zeros((cfg.levels, 8, B))
    for lvl in range(cfg.levels):
        N = level_resolution(cfg, lvl)
        s = x * N + 0.5
        g = np.floor(s)
        f = s - g
        for k in range(8):
            o = np.array([(k >> 0) & 1, (k >> 1) & 1, (k >> 2) & 1])
            w = np.prod(np.where(o[None, :] == 1, f, 1.0 - f), axis=1)
            idx = corner_index(cfg, lvl, g + o[None, :])
            idxs[lvl, k] = idx
            wts[lvl, k] = w
            feats[:, lvl * cfg.features:(lvl + 1) * cfg.features] += w[:, None] * tab[lvl, idx]
    return feats, idxs, wts


def init_params(cfg: Config, seed: int = 0):
    """Xavier-uniform weights and U(-1e-4, 1e-4) table, from a seeded generator."""
    rng = np.random.default_rng(seed)
    W = cfg.width
    shapes = [(W, cfg.levels * cfg.features + 6), (W, W), (W, W), (3, W)]
    mats = [rng.uniform(-1, 1, s) * np.sqrt(6.0 / (s[0] + s[1])) for s in shapes]
    table = rng.uniform(-1e-4, 1e-4, (cfg.levels, cfg.table, cfg.features))
    return [m.astype(np.float32) for m in mats], table.astype(np.float32)


def forward_backward(cfg: Config, mats, table, pos, normal, albedo, target=None, bf16_ops: bool = True):
    """One pass. Returns dict with y (B, 3) and, with a target, loss and the
    gradients dW (list like mats) and dtable (like table)."""
    rb = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    B = pos.shape[0]
    enc, idxs, wts = encode(cfg, table, pos, bf16_table=bf16_ops)
    x0 = np.concatenate([enc, np.asarray(normal, np.float64), np.asarray(albedo, np.float64)], axis=1)
    Ws = [rb(m) for m in mats]
    xs, zs = [rb(x0)], []
    h = xs[0]
    for i in range(3):
        z = h @ Ws[i].T
        zs.append(z)
        h = rb(np.maximum(z, 0.0))
        xs.append(h)
    y = h @ Ws[3].T
    out = {"y": y}
    if target is None:
        return out
    t = np.asarray(target, np.float64)
    out["loss"] = float(0.5 * np.sum((y - t) ** 2) / B)
    dz = rb((y - t) / B)
    dW = [None] * 4
    dW[3] = dz.T @ xs[3]
    for i in (2, 1, 0):
        dh = dz @ Ws[i + 1]
        dz = rb(dh * (zs[i] > 0.0))
        dW[i] = dz.T @ xs[i]
    dx0 = dz @ Ws[0]
    denc = dx0[:, :cfg.levels * cfg.features]
    dtab = np.zeros((cfg.levels, cfg.table, cfg.features))
    for lvl in range(cfg.levels):
        g = denc[:, lvl * cfg.features:(lvl + 1) * cfg.features]
        for k in range(8):
            np.add.at(dtab[lvl], idxs[lvl, k], wts[lvl, k][:, None] * g)
    out["dW"] = dW
    out["dtable"] = dtab
    return out


class Adam:
    """Adam over a list of float32 arrays (master copies)."""

    def __init__(self, cfg: Config, params):
        self.cfg = cfg
        self.m = [np.zeros_like(p, np.float64) for p in params]
        self.v = [np.zeros_like(p, np.float64) for p in params]
        self.t = 0

    def step(self, params, grads):
        c = self.cfg
        self.t += 1
        b1t = 1.0 - c.beta1 ** self.t
        b2t = 1.0 - c.beta2 ** self.t
        out = []
        for i, (p, g) in enumerate(zip(params, grads)):
            self.m[i] = c.beta1 * self.m[i] + (1.0 - c.beta1) * g
            self.v[i] = c.beta2 * self.v[i] + (1.0 - c.beta2) * g * g
            upd = c.lr * (self.m[i] / b1t) / (np.sqrt(self.v[i] / b2t) + c.eps)
            out.append((p.astype(np.float64) - upd).astype(np.float32))
        return out


def relative_rmse(est: np.ndarray, ref: np.ndarray) -> float:
    """sqrt(mean((est - ref)^2)) / mean(ref) over every voxel and channel."""
    est = np.asarray(est, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sqrt(np.mean((est - ref) ** 2)) / max(np.mean(ref), 1e-30))

"""Pins of the radiance-cache oracle (CPU): the encoding's closed forms, the
backward pass against central finite differences, and Adam's first step."""
from __future__ import annotations

import numpy as np

from oracle import radiance_cache as rc


def small_cfg(**kw):
    base = dict(levels=3, features=2, log2_table=8, n_min=2, n_max=16, width=8, extent=8.0)
    base.update(kw)
    return rc.Config(**base)


def test_level_resolutions():
    cfg = rc.Config()
    res = [rc.level_resolution(cfg, l) for l in range(cfg.levels)]
    assert res[0] == 16
    assert res[-1] in (255, 256)          # floor(16 * 16^(15/15)) up to rounding of b^15
    assert all(b > a for a, b in zip(res, res[1:]))
    b = 16.0 ** (1.0 / 15.0)
    assert all(abs(r - 16 * b ** l) < 1.0 for l, r in enumerate(res))


def test_corner_index_dense_and_hash():
    cfg = rc.Config()
    # level 0 (N = 16): dense, (16 + 2)^3 = 5832 <= 2^19
    assert rc.corner_index(cfg, 0, np.array([[1, 2, 3]]))[0] == 1 + 2 * 18 + 3 * 18 * 18
    # finest level: spatial hash, 32-bit products then mod 2^19
    top = cfg.levels - 1
    assert (rc.level_resolution(cfg, top) + 2) ** 3 > cfg.table
    c = np.array([[5, 7, 11]])
    h = (5 * 1) ^ ((7 * 2654435761) & 0xFFFFFFFF) ^ ((11 * 805459861) & 0xFFFFFFFF)
    assert rc.corner_index(cfg, top, c)[0] == h % (1 << 19)


def test_encoding_partition_of_unity_and_vertex_values():
    cfg = small_cfg()
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, cfg.extent, (50, 3))
    table = np.ones((cfg.levels, cfg.table, cfg.features), np.float32) * np.arange(1, cfg.levels + 1)[:, None, None]
    feats, idx, w = rc.encode(cfg, table, pos)
    assert np.allclose(w.sum(axis=1), 1.0)
    for l in range(cfg.levels):
        assert np.allclose(feats[:, 2 * l:2 * l + 2], l + 1)
    # a position on a grid vertex reads exactly that vertex's entry
    N = rc.level_resolution(cfg, 1)
    v = np.array([2, 3, 1])
    p = (v - 0.5) / N * cfg.extent
    t2 = rng.uniform(-1, 1, table.shape).astype(np.float32)
    f2, _, _ = rc.encode(cfg, t2, p[None, :])
    assert np.allclose(f2[0, 2:4], t2[1, rc.corner_index(cfg, 1, v[None, :])[0]], atol=1e-12)


def _loss(cfg, mats, table, batch):
    return rc.forward_backward(cfg, mats, table, *batch, bf16_ops=False)["loss"]


def test_backward_matches_finite_differences():
    cfg = small_cfg()
    rng = np.random.default_rng(7)
    B = 6
    mats = [rng.normal(0, 0.5, s) for s in [(cfg.width, cfg.levels * cfg.features + 6), (cfg.width, cfg.width), (cfg.width, cfg.width),
                                            (3, cfg.width)]]
    table = rng.normal(0, 0.5, (cfg.levels, cfg.table, cfg.features))
    batch = (rng.uniform(0, cfg.extent, (B, 3)), rng.uniform(-1, 1, (B, 3)), rng.uniform(0, 1, (B, 3)),
             rng.uniform(0, 1, (B, 3)))
    out = rc.forward_backward(cfg, mats, table, *batch, bf16_ops=False)
    eps = 1e-6
    for li, m in enumerate(mats):
        for _ in range(6):
            i, j = rng.integers(0, m.shape[0]), rng.integers(0, m.shape[1])
            mp = [x.copy() for x in mats]
            mm = [x.copy() for x in mats]
            mp[li][i, j] += eps
            mm[li][i, j] -= eps
            fd = (_loss(cfg, mp, table, batch) - _loss(cfg, mm, table, batch)) / (2 * eps)
            assert abs(fd - out["dW"][li][i, j]) <= 1e-6 + 1e-5 * abs(fd), (li, i, j, fd, out["dW"][li][i, j])
    _, idx, _ = rc.encode(cfg, table, batch[0])
    for lvl in range(cfg.levels):
        for e in np.unique(idx[lvl])[:4]:
            for f in range(cfg.features):
                tp, tm = table.copy(), table.copy()
                tp[lvl, e, f] += eps
                tm[lvl, e, f] -= eps
                fd = (_loss(cfg, mats, tp, batch) - _loss(cfg, mats, tm, batch)) / (2 * eps)
                assert abs(fd - out["dtable"][lvl, e, f]) <= 1e-6 + 1e-5 * abs(fd)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.0e-3, 65504.0], np.float32)
    r = rc.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                                   # tie to even
    assert r[2] == 1.0 + 2 ** -7                         # below the midpoint rounds down
    assert abs(r[3] + 3.0e-3) <= 3.0e-3 * 2 ** -8
    assert rc.bf16(np.array([1.0 + 3 * 2 ** -8], np.float32))[0] == 1.0 + 2 ** -6   # tie to even, upward


def test_adam_first_step_is_signed_lr():
    cfg = rc.Config()
    p = [np.array([1.0, -2.0, 0.5], np.float32)]
    g = [np.array([0.3, -1e-3, 0.0])]
    opt = rc.Adam(cfg, p)
    q = opt.step(p, g)[0]
    assert np.allclose(q, p[0] - cfg.lr * np.sign(g[0]), atol=1e-7)

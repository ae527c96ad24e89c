"""Pins of the oracle (CPU only): each function checked against what the
mathematics fixes -- closed forms, brute force on small inputs, invariants --
or against the reference's own constant table, so that a plausible slip in the
restatement (a dropped term, a wrong index, a swapped operand) fails here.
"""
from __future__ import annotations

import ctypes
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from doonengine_amd import scenes
from helpers import build_state, camera, cpu_chunk, materials_array

GOLDEN = Path(__file__).parent / "golden"
f32 = np.float32


@pytest.fixture(scope="module")
def o():
    return oracle.load("f32")


@pytest.fixture(scope="module")
def o64():
    return oracle.load("f64")


# ---------------------------------------------------------------------------
# codec (voxelShared.comp:103-118, 236-256)

def test_codec_constants_bits():
    # literal constants of the shader, as FP32 bit patterns
    assert np.array([0.00392156862], f32).view(np.uint32)[0] == 0x3B808080   # one ulp below 1/255
    assert np.array([1 / 255], f32).view(np.uint32)[0] == 0x3B808081
    assert np.array([0.0000152590219], f32).view(np.uint32)[0] == 0x37800080  # == float(1/65535)
    assert np.array([1 / 65535], f32).view(np.uint32)[0] == 0x37800080
    assert np.array([0.0001], f32).view(np.uint32)[0] == 0x38D1B717


def test_decompress_matches_closed_form(o, o64):
    rng = np.random.default_rng(1)
    k255, k65535 = f32(0.00392156862), f32(0.0000152590219)
    for _ in range(200):
        w = rng.integers(0, 2**32, 4, dtype=np.uint64).astype(np.uint32)
        got = o.decompress(w)
        b = lambda x, s: (int(x) >> s) & 255
        n = [b(w[0], 16), b(w[0], 8), b(w[0], 0)]
        exp_n = [(f32(v) * k255 - f32(0.5)) * f32(2.0) for v in n]
        alb = [f32(b(w[1], s)) * k255 for s in (24, 16, 8)]
        spec = [f32(b(w[1], 0)) * k255, f32(b(w[2], 24)) * k255, f32(b(w[2], 16)) * k255]
        dif = [f32(int(w[2]) & 0xFFFF) * k65535, f32(int(w[3]) >> 16) * k65535, f32(int(w[3]) & 0xFFFF) * k65535]
        exp = np.array(exp_n + alb + spec + dif + [int(w[0]) >> 24], f32)
        assert np.array_equal(got.astype(f32), exp)
        g64 = o64.decompress(w)
        assert np.allclose(g64, exp, rtol=1e-6, atol=1e-7)


def test_quantisation_round_trips():
    # stored codes survive decode -> encode: 8-bit with the shader literal, 16-bit with 1/65535
    v8 = np.arange(256, dtype=f32)
    assert np.array_equal(np.rint(v8 * f32(0.00392156862) * f32(255.0)), v8)
    v16 = np.arange(65536, dtype=f32)
    assert np.array_equal(np.rint(v16 * f32(0.0000152590219) * f32(65535.0)), v16)


# ---------------------------------------------------------------------------
# rank / select (voxelShared.comp:150-224)

def _chunk_from_mask(bits: np.ndarray) -> np.ndarray:
    words = np.packbits(bits.reshape(16, 32)[:, ::-1], axis=1).view(">u4").reshape(16).astype(np.uint32)
    c = np.zeros(24, np.uint32)
    c[7:23] = words
    cs = np.cumsum(bits)
    c[4:7] = [cs[127], cs[255], cs[383]]
    return c


def test_rank_select_brute_force(o):
    rng = np.random.default_rng(2)
    masks = [np.zeros(512, bool), np.ones(512, bool)]
    for p in (0.01, 0.1, 0.5, 0.9):
        for _ in range(40):
            masks.append(rng.random(512) < p)
    for pos in (0, 31, 32, 127, 128, 255, 256, 383, 384, 511):
        m = np.zeros(512, bool)
        m[pos] = True
        masks.append(m)
    for bits in masks:
        c = _chunk_from_mask(bits)
        set_pos = np.nonzero(bits)[0]
        for li in list(range(0, 512, 7)) + [127, 128, 255, 256, 383, 384, 511]:
            x, y, z = li % 8, (li // 8) % 8, li // 64
            assert o.rank(c, x, y, z) == int(bits[:li].sum())
        for n in range(0, len(set_pos) + 3):
            s = o.select(c, n)
            assert s == (int(set_pos[n]) if n < len(set_pos) else -1)
            if s >= 0:
                assert o.rank(c, s % 8, (s // 8) % 8, s // 64) == n   # select(rank) invariant


# ---------------------------------------------------------------------------
# chunk conversion (voxel.c:1391-1461)

def _brute_chunk(cv: np.ndarray, mats: np.ndarray):
    mat = cv[..., 0] >> 24
    solid = mat != 255
    opaque = solid & (mats["opacity"][np.minimum(mat, 255)] >= 1.0)
    pad = np.zeros((10, 10, 10), bool)
    pad[1:9, 1:9, 1:9] = opaque
    covered = (pad[2:, 1:9, 1:9] & pad[:-2, 1:9, 1:9] & pad[1:9, 2:, 1:9] & pad[1:9, :-2, 1:9] &
               pad[1:9, 1:9, 2:] & pad[1:9, 1:9, :-2])
    # out-of-chunk neighbours count as visible faces
    edge = np.zeros((8, 8, 8), bool)
    edge[[0, -1], :, :] = True
    edge[:, [0, -1], :] = True
    edge[:, :, [0, -1]] = True
    stored = solid & ~(covered & ~edge)
    bits = np.zeros(512, bool)
    order = []
    for z in range(8):
        for y in range(8):
            for x in range(8):
                if stored[x, y, z]:
                    bits[x + 8 * (y + 8 * z)] = True
                    order.append((x, y, z))
    return bits, order


def test_chunk_to_gpu_brute_force(o):
    rng = np.random.default_rng(3)
    mats = np.zeros(256, oracle.MATERIAL_DTYPE)
    mats["opacity"] = 1.0
    mats[2]["opacity"] = 0.5
    lut = np.array([int(f32(math.pow(float(f32(v) * f32(0.00392156862)), 2.2)) * f32(255.0)) for v in range(256)])
    for trial in range(60):
        p = [0.0, 1.0, 0.3, 0.7, 0.95][trial % 5]
        solid = rng.random((8, 8, 8)) < p
        mat = np.where(solid, rng.integers(0, 3, (8, 8, 8)), 255).astype(np.uint32)
        alb = rng.integers(0, 256, (8, 8, 8, 3)).astype(np.uint32)
        cv = np.zeros((8, 8, 8, 2), np.uint32)
        cv[..., 0] = (mat << 24) | 0x7F7F7F
        cv[..., 1] = (alb[..., 0] << 24) | (alb[..., 1] << 16) | (alb[..., 2] << 8)
        ch, vv = o.chunk_to_gpu(mats, cv)
        bits, order = _brute_chunk(cv, mats)
        assert np.array_equal(_chunk_from_mask(bits)[4:23], ch[4:23])
        assert len(vv) == len(order)
        for k, (x, y, z) in enumerate(order):
            a = alb[x, y, z]
            assert vv[k, 0] == cv[x, y, z, 0]
            exp = (lut[a[0]] << 24) | (lut[a[1]] << 16) | (lut[a[2]] << 8)
            assert vv[k, 1] == exp
            assert vv[k, 2] == 0 and vv[k, 3] == 0
        if p == 1.0:
            assert len(vv) == 8 ** 3 - 6 ** 3 or (mat == 2).any()


def test_full_solid_chunk_stores_shell(o):
    mats = np.zeros(256, oracle.MATERIAL_DTYPE)
    mats["opacity"] = 1.0
    cv = np.zeros((8, 8, 8, 2), np.uint32)
    cv[..., 0] = 0x007F7F7F
    ch, vv = o.chunk_to_gpu(mats, cv)
    assert len(vv) == 296     # 8^3 - 6^3: chunk borders always count as visible faces


# ---------------------------------------------------------------------------
# DDA (voxelShared.comp:300-475) vs exact traversal

def _exact_first_hit(solid: np.ndarray, org: np.ndarray, d: np.ndarray):
    """First solid voxel along the ray by exact slab intersection in FP64; returns
    (voxel xyz, entry t, margin): margin is the smaller of the gap to the next
    candidate's entry and the length of the ray inside the first candidate
    (a grazed edge or corner is a near-tie)."""
    idx = np.argwhere(solid)
    lo = idx.astype(np.float64)
    hi = lo + 1.0
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t1 = (lo - org) * inv
        t2 = (hi - org) * inv
    tn = np.nanmax(np.minimum(t1, t2), axis=1)
    tf = np.nanmin(np.maximum(t1, t2), axis=1)
    ok = (tf >= np.maximum(tn, 0.0))
    if not ok.any():
        return None, None, None
    tn = np.where(ok, np.maximum(tn, 0.0), np.inf)
    order = np.argsort(tn)
    gap = tn[order[1]] - tn[order[0]] if ok.sum() > 1 else np.inf
    span = tf[order[0]] - tn[order[0]]
    return tuple(idx[order[0]]), tn[order[0]], min(gap, span)


def test_dda_hits_match_exact_traversal(o):
    sc = scenes.config_small(seed=21, chunks=(3, 2, 3), occupancy=0.35, n_boxes=4)
    m, c, v, mats = build_state(sc)
    s = o.scene(sc.map_size, m, c, v, mats)
    solid = ((sc.normal_words >> 24) != 255).transpose(2, 1, 0)        # [x, y, z] voxels
    # voxel position of every pool slot
    pos_of = {}
    for t in np.nonzero(m[:, 0] == 2)[0]:
        cx, cy, cz = c[t, 0:3]
        bits = np.unpackbits(c[t, 7:23].astype(">u4").view(np.uint8)).reshape(16, 32)[:, ::-1].reshape(512)
        for k, li in enumerate(np.nonzero(bits)[0]):
            pos_of[int(m[t, 2]) + k] = (cx * 8 + li % 8, cy * 8 + (li // 8) % 8, cz * 8 + li // 64)
    rng = np.random.default_rng(4)
    n = 3000
    dims = np.array(solid.shape, np.float64)
    org = (rng.uniform(0.0, 1.0, (n, 3)) * dims).astype(f32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = (d + 1e-4).astype(f32)
    res, _ = o.cast_rays(s, org / f32(8.0), d)            # oracle works in chunk units
    agree, near_tie = 0, 0
    for i in range(n):
        o64 = org[i].astype(np.float64)
        if solid[tuple(np.floor(o64).astype(int).clip(0, dims.astype(int) - 1))]:
            continue                                     # origin inside a solid voxel: the DDA hits it at once
        exp, t, gap = _exact_first_hit(solid, o64, d[i].astype(np.float64))
        got = pos_of.get(int(res["voxelIndex"][i])) if res["hit"][i] else None
        if exp is None:
            assert got is None or True                   # grazing exits are near-ties by construction
            agree += got is None
            continue
        if got == tuple(int(a) for a in exp):
            agree += 1
        else:
            assert gap < 2e-3 or t < 2e-3, (i, got, exp, t, gap)
            near_tie += 1
    assert agree > 0.95 * (agree + near_tie)


def test_dda_step_counts_closed_form(o):
    # a map whose chunks are all loaded but hold one voxel in a far corner
    X, Y, Z = 5, 4, 6
    T = X * Y * Z
    m = np.zeros((T, 3), np.uint32)
    c = np.zeros((T, 24), np.uint32)
    m[:, 0] = 2
    li = 511                                             # voxel (7, 7, 7)
    c[:, 7 + li // 32] = 1 << (li % 32)
    c[:, 4:7] = 0
    v = np.zeros((1, 4), np.uint32)
    v[0, 0] = 0x007F7F7F
    mats = np.zeros(256, oracle.MATERIAL_DTYPE)
    mats["opacity"] = 1.0
    s = o.scene((X, Y, Z), m, c, v, mats)
    rng = np.random.default_rng(6)
    checked = 0
    for _ in range(400):
        org = rng.uniform(0.05, 0.95, 3) * [X, Y, Z]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        d = (d + 1e-4).astype(f32)
        org = org.astype(f32)
        res, _ = o.cast_rays(s, org[None], d[None], ignore_first=False)
        if res["hit"][0]:
            continue
        # chunk-grid trips = cells visited before leaving the map = 1 + crossings (exact, FP64)
        o64, d64 = org.astype(np.float64), d.astype(np.float64)
        tmax = min((([X, Y, Z][k] if d64[k] > 0 else 0) - o64[k]) / d64[k] for k in range(3))
        crossings = 0
        times = []
        for k in range(3):
            a = np.floor(o64[k])
            b = np.floor(o64[k] + d64[k] * tmax * (1 - 1e-12))
            crossings += int(abs(b - a))
            step = 1 if d64[k] > 0 else -1
            first = (a + (1 if step > 0 else 0) - o64[k]) / d64[k]
            times += [first + j / abs(d64[k]) for j in range(int(abs(b - a)))]
        times = np.sort(times)
        if len(times) > 1 and np.min(np.diff(times)) < 1e-4:
            continue                                     # near-simultaneous crossings step together
        assert res["mapSteps"][0] == 1 + crossings
        # voxel trips: each loaded chunk entered is crossed voxel by voxel (8 per chunk unit)
        checked += 1
    assert checked > 200


def test_fp64_oracle_agrees_on_draw_hits(o, o64):
    sc = scenes.config_small()
    m, c, v, mats = build_state(sc)
    _, _, iv, ivp = camera(sc)
    s32 = o.scene(sc.map_size, m, c, v, mats)
    s64 = o64.scene(sc.map_size, m, c, v, mats)
    _, p32, _ = o.draw(s32, iv, ivp, sc.width, sc.height)
    _, p64, _ = o64.draw(s64, iv, ivp, sc.width, sc.height)
    assert (p32["voxelIndex"] == p64["voxelIndex"]).mean() > 0.99


# ---------------------------------------------------------------------------
# draw (voxelDraw.comp) on closed-form scenes

def _sky_closed_form(dirs, bot, top):
    a = dirs[..., 1:2] * 0.5 + 1.0
    return np.asarray(bot) * (1 - a) + np.asarray(top) * a


def test_draw_empty_map_is_sky(o):
    sc = scenes.config_small()
    T = int(np.prod(sc.map_size))
    m = np.zeros((T, 3), np.uint32)
    c = np.zeros((T, 24), np.uint32)
    v = np.zeros((1, 4), np.uint32)
    mats = materials_array(sc)
    s = o.scene(sc.map_size, m, c, v, mats)
    _, _, iv, ivp = camera(sc)
    rgba, px, cnt = o.draw(s, iv, ivp, sc.width, sc.height)
    assert (px["voxelIndex"] == 0xFFFFFFFF).all() and cnt["hits"] == 0
    # closed-form ray directions: normalize(M * (ndc, 0, 1)).xyz + eps
    W, H = sc.width, sc.height
    xs, ys = np.meshgrid(np.arange(W) / W * 2 - 1, np.arange(H) / H * 2 - 1)
    M = ivp.astype(np.float64)
    q = xs[..., None] * M[0, :3] + ys[..., None] * M[1, :3] + M[3, :3]
    d = q / np.linalg.norm(q, axis=-1, keepdims=True) + 1e-4
    exp = np.power(np.power(_sky_closed_form(d, (0.71, 0.85, 0.90), (0.0, 0.45, 0.74)), 2.2), 0.4545)
    np.testing.assert_allclose(rgba[..., :3], exp, rtol=2e-5, atol=2e-6)
    assert (rgba[..., 3] == -1.0).all()


def test_draw_view_modes_voxel_layer(o):
    """A layer of identical voxels (y = 3 of a 1x1x1 map) seen from above: every
    view mode of voxel_color (voxelDraw.comp:32-61) has a closed form."""
    m = np.array([[2, 0, 0]], np.uint32)
    bits = np.zeros(512, bool)
    bits[[x + 8 * (3 + 8 * z) for z in range(8) for x in range(8)]] = True
    c = _chunk_from_mask(bits)[None].copy()
    word = np.array([(0 << 24) | (127 << 16) | (255 << 8) | 127, (200 << 24) | (100 << 16) | (50 << 8) | 64,
                     (32 << 24) | (16 << 16) | 0x8000, (0x4000 << 16) | 0xC000], np.uint32)
    v = np.tile(word, (64, 1))
    mats = np.zeros(256, oracle.MATERIAL_DTYPE)
    mats["opacity"] = 1.0
    mats[0]["specular"] = 0.25
    s = o.scene((1, 1, 1), m, c, v, mats)
    from doonengine_amd.scenes import Scene
    sc = Scene("layer", (1, 1, 1), np.zeros((8, 8, 8), np.uint32), np.zeros((8, 8, 8), np.uint32), [],
               cam_pos=(0.5, 2.0, 0.5), cam_orient=(89.0, 0.0, 0.0), width=16, height=16)
    _, _, iv, ivp = camera(sc)
    vox = o.decompress(word)
    alb, spec, dif = vox[3:6].astype(np.float64), vox[6:9].astype(np.float64), vox[9:12].astype(np.float64)
    g = 0.4545
    exp = {0: np.power(dif * 0.75 * alb + spec * 0.25, g), 1: np.power(alb, g), 2: np.power(dif * 0.75, g),
           3: np.power(spec * 0.25, g), 4: np.power(np.abs(vox[0:3].astype(np.float64)), g),
           5: np.array([0.0, 1.0, 0.0])}
    for mode in range(6):
        rgba, px, _ = o.draw(s, iv, ivp, 16, 16, view_mode=mode)
        top = (px["voxelIndex"] != 0xFFFFFFFF) & (np.abs(px["face"][..., 1]) == 1.0)
        assert top.sum() > 20
        np.testing.assert_allclose(rgba[top][:, :3], np.broadcast_to(exp[mode], (top.sum(), 3)), rtol=1e-5,
                                   atol=1e-6)


# ---------------------------------------------------------------------------
# lighting (voxelLighting.comp) on closed-form scenes

def _floor_scene(sx=4, sz=4, material_specular=0.0, reflect_type=0, shininess=0, emissive_walls=False):
    """A one-voxel-thick opaque floor at voxel y = 0 across an sx x 2 x sz chunk map."""
    X, Y, Z = sx, 2, sz
    nw = np.full((Z * 8, Y * 8, X * 8), 0xFFFFFFFF, np.uint32)
    aw = np.zeros_like(nw)
    nw[:, 0, :] = (0 << 24) | (127 << 16) | (255 << 8) | 127     # normal (0, 1, 0)
    aw[:, 0, :] = (255 << 24) | (255 << 16) | (255 << 8)
    mats = [scenes.Material(albedo=(255, 255, 255), specular=material_specular, shininess=shininess,
                            reflect_type=reflect_type)]
    return scenes.Scene("floor", (X, Y, Z), nw, aw, mats, cam_pos=(2, 1.5, 2), cam_orient=(90, 0, 0), width=8,
                        height=8)


def _all_requests(o, m, c):
    mm = m.copy()
    mm[(mm[:, 0] & 3) == 2, 0] |= 4
    return o.visible_requests(mm, c), mm


def test_lighting_open_floor_first_sample(o, o64):
    sc = _floor_scene()
    m, c, v, mats = build_state(sc)
    req, mm = _all_requests(o, m, c)
    sun = np.array([0.3, 1.0, 0.2], f32)
    sun = sun / np.float32(np.linalg.norm(sun))
    strength = 0.2
    for orc in (o, o64):
        s = orc.scene(sc.map_size, mm, c, v, mats, sun=(strength,) * 3)
        vo, mo, co, preq, voxof, cnt = orc.lighting(s, req, time=1.0, sun_dir=sun, spec_points=np.zeros((15, 3), f32),
                                                     ambient=(0.01, 0.01, 0.01))
        valid = voxof != 0xFFFFFFFF
        # top-face voxels away from chunk borders see the open sky: the first-sample
        # diffuse ray goes straight up and escapes (sun-cosine term), the shadow ray
        # toward the sun escapes (full sun, no cosine: voxelLighting.comp:78-79)
        # normal bytes (127, 255, 127) decode to (-0.0039, 1, -0.0039)
        up = orc.decompress(v[0])[0:3].astype(np.float64) + 1e-4
        up = up / np.linalg.norm(up) + 1e-4
        exp = 0.01 + strength * max(float(np.dot(up, sun)), 0.0) + strength
        np.testing.assert_allclose(preq[valid, 0], exp, rtol=1e-5)
        np.testing.assert_allclose(preq[valid, 3:6], 0.0)
        assert (co[mo[:, 0] & 3 == 2, 3] == 1).all()           # one sample taken
        assert ((mo[:, 0] & 4) == 0).all()                     # visible cleared


def test_lighting_fixed_direct_rays_unoccluded(o):
    sc = _floor_scene()
    m, c, v, mats = build_state(sc)
    req, mm = _all_requests(o, m, c)
    s = o.scene(sc.map_size, mm, c, v, mats, sun=(0.2, 0.2, 0.2))
    pts = oracle.golden_points(16).astype(f32)
    sun = np.array([0.0, 1.0, 0.0], f32)
    _, _, _, preq, voxof, cnt = o.lighting(s, req, time=1.0, sun_dir=sun, spec_points=np.zeros((15, 3), f32),
                                           direct_points=pts, ambient=(0.0, 0.0, 0.0))
    valid = voxof != 0xFFFFFFFF
    # all 16 shadow rays escape: direct term = 16 * sun / 16; plus the diffuse escape cosine
    up = o.decompress(v[0])[0:3].astype(np.float64) + 1e-4
    up = up / np.linalg.norm(up) + 1e-4
    exp = 0.2 + 0.2 * up[1]
    np.testing.assert_allclose(preq[valid, 0], exp, rtol=1e-5)
    assert cnt["calls"] == valid.sum() * (16 + 1)


def test_lighting_mirror_floor_specular_closed_form(o):
    sc = _floor_scene(material_specular=1.0, reflect_type=1, shininess=100)
    m, c, v, mats = build_state(sc)
    req, mm = _all_requests(o, m, c)
    s = o.scene(sc.map_size, mm, c, v, mats, sun=(0.5, 0.5, 0.5))
    pts = np.array(json.loads((GOLDEN / "sphere_points_15.json").read_text())["points"], f32)
    sun = np.array([0.0, 1.0, 0.0], f32)
    cam = (2.0, 1.5, 2.0)
    _, _, _, preq, voxof, _ = o.lighting(s, req, time=1.0, sun_dir=sun, spec_points=pts, cam_pos=cam)
    valid = np.nonzero(voxof != 0xFFFFFFFF)[0]
    checked = 0
    for slot in valid[::37]:
        r, lane = divmod(slot, 32)
        t, blk = int(req[r]) >> 4, int(req[r]) & 15
        sel = o.select(c[t], lane + blk * 32)
        cp = np.array([sel % 8, (sel // 8) % 8, sel // 64], np.float64)
        if cp[1] != 0:
            continue
        p = cp * 0.125 + c[t, 0:3] + 0.0625
        n = o.decompress(v[0])[0:3].astype(np.float64)          # (127, 255, 127) bytes, slightly tilted
        p = p + (0.0625 - 1e-4) * n
        view = p - np.array(cam)
        refl = view / np.linalg.norm(view)
        refl = refl - 2 * np.dot(n, refl) * n
        acc = np.zeros(3)
        for q in pts.astype(np.float64):
            dd = refl * 100 + q
            dd = dd / np.linalg.norm(dd) + 1e-4
            if np.dot(dd, sun) > 0.99:
                acc += 0.5                                      # sun highlight (no albedo multiplier)
            else:
                a = dd[1] * 0.5 + 1.0                            # reflectType 1: sky gradient
                acc += (np.array([0.71, 0.85, 0.90]) * (1 - a) + np.array([0.0, 0.45, 0.74]) * a) * (254 / 255)
        np.testing.assert_allclose(preq[slot, 3:6], acc / 15, rtol=2e-4)
        checked += 1
    assert checked > 5


def _shell_scene():
    """One chunk whose border voxels are emissive (material 1) around a single
    diffuse voxel (material 0) at (4, 1, 4): every ray from it hits emission."""
    nw = np.full((8, 8, 8), 0xFFFFFFFF, np.uint32)
    aw = np.zeros_like(nw)
    zz, yy, xx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    shell = (xx == 0) | (xx == 7) | (yy == 0) | (yy == 7) | (zz == 0) | (zz == 7)
    nw[shell] = (1 << 24) | (127 << 16) | (255 << 8) | 127
    aw[shell] = (180 << 24) | (120 << 16) | (60 << 8)
    nw[4, 1, 4] = (0 << 24) | (127 << 16) | (255 << 8) | 127
    aw[4, 1, 4] = (200 << 24) | (200 << 16) | (200 << 8)
    mats = [scenes.Material(albedo=(200, 200, 200)), scenes.Material(albedo=(180, 120, 60), emissive=True)]
    return scenes.Scene("shell", (1, 1, 1), nw, aw, mats, cam_pos=(0.5, 0.5, 0.5), cam_orient=(0, 0, 0), width=4,
                        height=4)


def test_lighting_running_mean_closed_form(o):
    """stored = (old * min(n, maxSamples) + sum_i (ambient + emission)) / (min(n, max) + numSamples)
    (voxelLighting.comp:224, 256-263), exactly in FP32, for first and later samples."""
    sc = _shell_scene()
    m, c, v, mats = build_state(sc)
    req, mm = _all_requests(o, m, c)
    amb = f32(0.01)
    t = int(np.nonzero((m[:, 0] & 3) == 2)[0][0])
    for n_prev, max_s, nds, old in ((0, 1000, 1, 0), (7, 4, 2, 30000), (3, 1000, 3, 12345), (50, 20, 1, 65535)):
        cc = c.copy()
        cc[t, 3] = n_prev
        vv = v.copy()
        vv[:, 2] = (vv[:, 2] & 0xFFFF0000) | old
        vv[:, 3] = (old << 16) | old
        s = o.scene(sc.map_size, mm, cc, vv, mats, sun=(0.6, 0.6, 0.6))
        _, _, co, preq, voxof, _ = o.lighting(s, req, time=77.0, sun_dir=np.array([0, 1, 0], f32),
                                              spec_points=np.zeros((15, 3), f32), num_diffuse=nds,
                                              max_diffuse=max_s, ambient=(amb, amb, amb))
        # emission = the shell's stored albedo, which went through the chunk conversion's gamma LUT
        lin = np.array([int(f32(math.pow(float(f32(b) * f32(0.00392156862)), 2.2)) * f32(255.0))
                        for b in (180, 120, 60)], f32) * f32(0.00392156862)
        acc = np.zeros(3, f32)
        for _ in range(nds):
            acc = acc + amb
            acc = acc + lin
        k = f32(min(n_prev, max_s))
        oldf = f32(old) * f32(0.0000152590219)
        exp = (oldf * k + acc) / (k + f32(nds))
        # the diffuse voxel is the one whose normal word has material 0
        slots = [i for i in np.nonzero(voxof != 0xFFFFFFFF)[0] if (vv[voxof[i], 0] >> 24) == 0]
        assert len(slots) == 1
        assert np.array_equal(preq[slots[0], :3].astype(f32), exp), (n_prev, max_s, nds, preq[slots[0], :3], exp)
        assert co[t, 3] == n_prev + 1


def test_rng_closed_form(o):
    rng = np.random.default_rng(7)
    seeds = rng.uniform(0, 3000, 2000).astype(f32)
    exact = 0
    for s in seeds:
        g = o.rand(float(s))
        assert -1.0 <= g < 1.0
        # fract(sin(s) * 43758.5453) * 2 - 1 in FP32, with sin(s) the correctly rounded
        # value or one of its neighbours (the C library's sinf is within 1 ulp)
        sn = f32(math.sin(float(s)))
        cands = []
        for sv in (sn, np.nextafter(sn, f32(2)), np.nextafter(sn, f32(-2))):
            v = f32(sv) * f32(43758.5453)
            cands.append((v - np.floor(v)) * f32(2.0) - f32(1.0))
        assert g in cands
        exact += g == cands[0]
    assert exact >= 0.95 * len(seeds)
    for s in seeds[:300]:
        p = o.rand_unit_sphere(float(s))
        assert float(np.dot(p.astype(np.float64), p.astype(np.float64))) < 1.0
        first = np.array([o.rand(float(s)), o.rand(float(s * f32(2.0))), o.rand(float(s * f32(3.0)))], f32)
        if float(np.dot(first, first)) < 1.0:
            assert np.array_equal(p, first)


# ---------------------------------------------------------------------------
# visible-chunk requests (voxel.c:719-762, 1463-1489)

def test_visible_requests_brute_force(o):
    rng = np.random.default_rng(8)
    T = 700
    m = np.zeros((T, 3), np.uint32)
    m[:, 0] = rng.integers(0, 4, T) | (rng.integers(0, 2, T) << 2)
    c = np.zeros((T, 24), np.uint32)
    counts = rng.integers(1, 513, T)
    for t in range(T):
        bits = np.zeros(512, bool)
        bits[rng.choice(512, counts[t], replace=False)] = True
        c[t] = _chunk_from_mask(bits)
    fresh = (rng.random(T) < 0.1).astype(np.uint8)
    for split, frame in ((1, 0), (3, 1), (5, 4)):
        got = o.visible_requests(m, c, fresh=fresh, split=split, frame=frame)
        exp = []
        for t in range(T):
            f = int(m[t, 0])
            if f & 3 != 2 or not f & 4:
                continue
            if t % split != frame and not fresh[t]:
                continue
            exp += [(t << 4) | b for b in range(int(math.ceil(counts[t] / 32)))]
        assert np.array_equal(got, np.array(exp, np.uint32))


# ---------------------------------------------------------------------------
# the reference's constant table

def test_sphere_points_table_is_near_golden_angle():
    pts = np.array(json.loads((GOLDEN / "sphere_points_15.json").read_text())["points"])
    assert pts.shape == (15, 3)
    np.testing.assert_allclose(pts, oracle.golden_points(15), atol=1.5e-6)
    np.testing.assert_allclose(np.linalg.norm(pts, axis=1), 1.0, atol=3e-6)


# ---------------------------------------------------------------------------
# host-thread splits of the oracle's draw and lighting (used for the full-size
# parity checks) give the single-thread result exactly

def test_threaded_oracle_equals_single_thread(o, o64):
    sc = scenes.config_small(seed=12)
    sc.materials[1].specular = 0.7                    # specular rays: visible-bit propagation across ranges
    sc.materials[1].shininess = 3
    m, c, v, mats = build_state(sc)
    for k in range(4):
        mats[k]["specular"] = sc.materials[k].specular
        mats[k]["shininess"] = sc.materials[k].shininess
    _, _, iv, ivp = camera(sc)
    for orc in (o, o64):
        s = orc.scene(sc.map_size, m, c, v, mats)
        f1 = m[:, 0].copy()
        r1, p1, c1 = orc.draw(s, iv, ivp, sc.width, sc.height, flags_out=f1)
        for th in (3, 7):
            fk = m[:, 0].copy()
            rk, pk, ck = orc.draw(s, iv, ivp, sc.width, sc.height, flags_out=fk, threads=th)
            assert np.array_equal(r1.view(np.uint32), rk.view(np.uint32)) and np.array_equal(p1, pk)
            assert np.array_equal(f1, fk)
            assert {k: ck[k] for k in ("calls", "sMap", "sVox", "hits")} == \
                {k: c1[k] for k in ("calls", "sMap", "sVox", "hits")}
        mm = m.copy()
        mm[:, 0] = f1
        req = orc.visible_requests(mm, c)
        assert len(req) > 20
        s2 = orc.scene(sc.map_size, mm, c, v, mats)
        kw = dict(time=321.0, sun_dir=np.array([-0.57735, 0.57735, -0.57735], np.float32),
                  spec_points=oracle.golden_points(15).astype(np.float32), cam_pos=sc.cam_pos)
        base = orc.lighting(s2, req, **kw)
        assert (base[1][:, 0] != mm[:, 0]).any()          # some visible bits change
        for th in (2, 5, len(req) + 3):
            got = orc.lighting(s2, req, threads=th, **kw)
            for a, b in zip(base[:5], got[:5]):
                assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
            assert {k: got[5][k] for k in ("calls", "sMap", "sVox", "hits")} == \
                {k: base[5][k] for k in ("calls", "sMap", "sVox", "hits")}


def _column_tops_brute(present: np.ndarray) -> np.ndarray:
    """[q][z][x]: the highest present tile's y over the columns (x', z') with
    x' >= x when q & 1 else x' <= x, and z' >= z when q & 2 else z' <= z."""
    Z, Y, X = present.shape
    ys = np.arange(Y)[None, :, None]
    top = np.where(present, ys, -1).max(axis=1) if Y else np.full((Z, X), -1)
    out = np.full((4, Z, X), -1, np.int64)
    for q in range(4):
        for z in range(Z):
            for x in range(X):
                xs = slice(x, X) if q & 1 else slice(0, x + 1)
                zs = slice(z, Z) if q & 2 else slice(0, z + 1)
                out[q, z, x] = top[zs, xs].max()
    return out


@pytest.mark.parametrize("shape,fill", [((5, 3, 7), 0.15), ((1, 4, 1), 0.5), ((4, 2, 3), 0.0), ((6, 5, 4), 0.02)])
def test_present_column_tops_brute_force(o, o64, shape, fill):
    """The column tables of the lighting walk's second stop rule, against a
    brute-force maximum over each cell's reachable columns (present = map
    flags & 3 != 0, loaded or not)."""
    Z, Y, X = shape
    rng = np.random.default_rng(Z * 100 + Y * 10 + X)
    flags = np.where(rng.random((Z, Y, X)) < fill, rng.integers(1, 4, (Z, Y, X)), 0).astype(np.uint32)
    map_ = np.zeros((X * Y * Z, 3), np.uint32)
    map_[:, 0] = flags.reshape(-1)                       # tile index x + X (y + Y z)
    want = _column_tops_brute(flags != 0)
    for orc in (o, o64):
        s = orc.scene((X, Y, Z), map_, np.zeros((X * Y * Z, 24), np.uint32), np.zeros((1, 4), np.uint32),
                      np.zeros((256, 8), np.uint32))
        got = np.zeros(4 * X * Z, np.int32)
        orc.lib.or_present_column_tops(ctypes.byref(s), got.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(got.reshape(4, Z, X), want)


def test_stop_outside_present_box_changes_only_counters(o, o64):
    """Lighting rays end once they are beyond the box of present tiles on the
    side they move toward, and rays that do not move down once above every
    present tile of the columns they can still reach (the rules of the GPU's
    lighting walk, which the oracle applies while or_stop_outside is set).
    With the rules on and off the
    lighting pass gives the same stored words, pre-quantisation values, flags
    (visible bits and the requests of unloaded tiles), lastUsed stamps and
    sample counts, and the same rays, voxel trips and hits; only the
    chunk-grid trips drop."""
    sc = scenes.config_small(seed=17, chunks=(6, 5, 6), occupancy=0.2)
    for k in range(4):
        sc.materials[k].specular = 0.5 if k == 1 else 0.0
        sc.materials[k].shininess = 4
    m, c, v, mats = build_state(sc)
    mats[1]["specular"], mats[1]["shininess"] = 0.5, 4
    loaded = np.nonzero((m[:, 0] & 3) == 2)[0]
    req, mm = _all_requests(o, m, c)
    mm[loaded[::5], 0] = (mm[loaded[::5], 0] & ~np.uint32(7)) | 1     # present, not loaded: rays request them
    req = o.visible_requests(mm, c)
    assert len(req) > 20
    kw = dict(time=321.0, sun_dir=np.array([-0.57735, 0.57735, -0.57735], np.float32), num_diffuse=2,
              spec_points=oracle.golden_points(15).astype(np.float32), direct_points=oracle.golden_points(8),
              cam_pos=sc.cam_pos)
    for orc in (o, o64):
        flag = ctypes.c_int.in_dll(orc.lib, "or_stop_outside")
        s = orc.scene(sc.map_size, mm, c, v, mats)
        try:
            flag.value = 0
            full = orc.lighting(s, req, **kw)
            flag.value = 1
            stop = orc.lighting(s, req, **kw)
        finally:
            flag.value = 1
        for a, b in zip(full[:5], stop[:5]):
            assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
        assert ((stop[1][:, 0] & 3) == 3).any()                         # some requests were made
        assert {k: stop[5][k] for k in ("calls", "sVox", "hits")} == {k: full[5][k] for k in ("calls", "sVox", "hits")}
        assert stop[5]["sMap"] < 0.9 * full[5]["sMap"]

"""The oracle against golden vectors produced by the reference's own GLSL shaders
(assets/shaders/*.comp on Mesa llvmpipe; tests/golden/make_llvmpipe_golden.py).

llvmpipe's float rounding is its own (it may fuse multiply-adds; its sin,
normalize and round are implementation-defined), so the bar is: DDA hit voxel,
chunk-grid and voxel trip counts identical on every pixel; colours within 1e-6;
stored lighting codes identical on >= 99.9% of voxels and never more than 1 LSB
apart; pre-quantisation values within 1e-4 relative on >= 99.9%.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).parent / "golden"


def _mats(u32):
    return np.ascontiguousarray(u32, np.uint32).view(oracle.MATERIAL_DTYPE).reshape(256)


@pytest.fixture(scope="module")
def draw_fx():
    return np.load(GOLDEN / "llvmpipe_draw.npz")


@pytest.fixture(scope="module")
def light_fx():
    return np.load(GOLDEN / "llvmpipe_light.npz")


@pytest.mark.parametrize("tag,mode", [("opaque", 0), ("opaque", 5), ("transparent", 0)])
def test_draw_matches_reference_shader(draw_fx, tag, mode):
    o = oracle.load("f32")
    f = draw_fx
    s = o.scene(tuple(f[f"{tag}_mapsize"]), f[f"{tag}_map"], f[f"{tag}_chunks"], f[f"{tag}_voxels"],
                _mats(f[f"{tag}_mats"]))
    rgba_ref = f[f"{tag}_{mode}_rgba"]
    H, W = rgba_ref.shape[:2]
    flags = f[f"{tag}_map"][:, 0].copy()
    rgba, px, _ = o.draw(s, f[f"{tag}_iv"], f[f"{tag}_ivp"], W, H, view_mode=mode, flags_out=flags)
    dbg = f[f"{tag}_{mode}_dbg"]
    assert (px["voxelIndex"] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(px["voxelIndex"], dbg[..., 0])
    assert np.array_equal(px["mapSteps"], dbg[..., 1])
    assert np.array_equal(px["voxelSteps"], dbg[..., 2])
    assert np.array_equal(flags, f[f"{tag}_{mode}_flags"])
    inside = rgba[..., 3] != -1.0          # depth is undefined where the map box is missed (Q14)
    np.testing.assert_allclose(rgba[..., :3], rgba_ref[..., :3], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(rgba[inside, 3], rgba_ref[inside, 3], rtol=1e-6)


@pytest.mark.parametrize("tag", ["first", "rand1", "rand5", "spec_first", "spec_rand"])
def test_lighting_matches_reference_shader(light_fx, tag):
    o = oracle.load("f32")
    f = light_fx
    tm, nds, maxd, cx, cy, cz = f[f"{tag}_params"]
    s = o.scene(tuple(f[f"{tag}_mapsize"]), f[f"{tag}_map"], f[f"{tag}_chunks"], f[f"{tag}_voxels"],
                _mats(f[f"{tag}_mats"]))
    req = f[f"{tag}_req"]
    # the hash values llvmpipe's sin produced for every seed of this frame
    o.set_rand_override(f[f"{tag}_rand_seeds"], f[f"{tag}_rand_vals"])
    try:
        vo, mo, co, preq, voxof, _ = o.lighting(s, req, time=float(tm), num_diffuse=int(nds), max_diffuse=int(maxd),
                                                sun_dir=f["sun"], spec_points=f["points"], cam_pos=(cx, cy, cz))
    finally:
        o.set_rand_override(None)
    idx = f[f"{tag}_idx"]
    codes = f[f"{tag}_codes"]
    mine = vo[idx]
    assert len(idx) > 1000
    exact = (mine == codes).all(axis=1)
    assert exact.mean() >= 0.999, exact.mean()
    # stored channels: spec 8-bit x3, diffuse 16-bit x3
    def channels(w):
        return np.stack([w[:, 1] & 255, w[:, 2] >> 24, (w[:, 2] >> 16) & 255, w[:, 2] & 0xFFFF, w[:, 3] >> 16,
                         w[:, 3] & 0xFFFF], axis=1).astype(np.int64)
    d = np.abs(channels(mine) - channels(codes))
    assert (d <= 1).mean() >= 0.999
    assert np.array_equal(mine[:, 0], codes[:, 0]) and np.array_equal(mine[:, 1] >> 8, codes[:, 1] >> 8)
    # pre-quantisation values, matched by pool index
    valid = voxof != 0xFFFFFFFF
    pos = {int(v): k for k, v in enumerate(idx)}
    ref_vals = f[f"{tag}_prequant"][[pos[int(v)] for v in voxof[valid]]]
    rel = np.abs(preq[valid] - ref_vals) / np.maximum(np.abs(ref_vals), 1e-6)
    assert (rel.max(axis=1) < 1e-4).mean() >= 0.999
    # visible propagation by specular rays (the harness keeps the bit set on lit chunks)
    lit = np.zeros(len(mo), bool)
    lit[np.unique(req >> 4)] = True
    assert np.array_equal(mo[~lit, 0], f[f"{tag}_flags_after"][~lit])

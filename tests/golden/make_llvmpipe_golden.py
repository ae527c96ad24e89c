"""Writes golden vectors produced by the reference's own GLSL shaders.

Runs assets/shaders/voxelDraw.comp and voxelLighting.comp (read at run time from
the reference tree by oracle/_ref/libref_llvmpipe.so) on Mesa llvmpipe, on small
seeded scenes, and stores inputs and outputs as compressed npz fixtures next to
this script:

  llvmpipe_draw.npz      draw cast: RGBA, map flags after, per-pixel (hit voxel
                         index, chunk-grid trips, voxel trips)
  llvmpipe_light.npz     lighting: first-sample and random-sample frames, a
                         scene with specular materials; stored voxel words and
                         pre-quantisation values of the lit voxels

The harness-local shader variant only adds counters / debug outputs and moves
the racy in-shader writes (sample-count increment, visible clear, in-place
lighting store) out, as described in oracle/ref_llvmpipe.c.
Run: python tests/golden/make_llvmpipe_golden.py (needs the reference tree).
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
from oracle.ref_llvmpipe import Llvmpipe  # noqa: E402
from doonengine_amd import scenes  # noqa: E402
from helpers import build_state, camera  # noqa: E402

SUN = np.array([-1.0, 1.0, -1.0], np.float32) * (np.float32(1.0) / np.sqrt(np.float32(3.0)))
POINTS = np.array(json.loads((HERE / "sphere_points_15.json").read_text())["points"], np.float32)


def mats_u32(mats):
    return np.ascontiguousarray(mats).view(np.uint32).reshape(256, 8)


def frame_seeds(time: float, nds: int, dbl: int, trials: int = 48) -> np.ndarray:
    """Every seed the lighting shader can hash in one frame (voxelLighting.comp:
    47-57, 159-168, 259-260), computed with the shader's FP32 seed arithmetic;
    rejection loops are covered for `trials` attempts."""
    f = np.float32
    t = f(time)
    out = []

    def rus(base):
        s = f(base)
        for _ in range(trials):
            out.extend([s, s * f(2.0), s * f(3.0)])
            s = s + f(1.0)

    for i in range(nds):
        seed = t * f(i + 1)
        for j in range(dbl):
            rus(seed + f(j))
            out.append((seed + f(dbl)) + f(j))
        rus(t * f(i + 1 + nds))
    for j in range(dbl):
        rus(t + f(j))
    return np.unique(np.array(out, np.float32))


def main():
    ref = Llvmpipe(instrument=True)
    o = oracle.load("f32")
    out = {"renderer": ref.renderer}

    # ---- draw: opaque scene and a transparent/refractive material scene
    draw = {}
    for tag, sc in (("opaque", scenes.config_small()), ("transparent", scenes.config_small(seed=13))):
        if tag == "transparent":
            sc.materials[0].opacity = 0.4
        m, c, v, mats = build_state(sc)
        if tag == "transparent":
            mats[0]["refractIndex"] = 1.3
        W, H = 64, 48
        _, _, iv, ivp = camera(sc, W, H)
        s = ref.scene(sc.map_size, m, c, v, mats)
        for mode in ((0, 5) if tag == "opaque" else (0,)):
            rgba, mo, dbg, _ = ref.draw(s, iv, ivp, W, H, view_mode=mode)
            draw[f"{tag}_{mode}_rgba"] = rgba
            draw[f"{tag}_{mode}_dbg"] = dbg
            draw[f"{tag}_{mode}_flags"] = mo[:, 0].copy()
        draw[f"{tag}_map"], draw[f"{tag}_chunks"], draw[f"{tag}_voxels"] = m, c, v
        draw[f"{tag}_mats"] = mats_u32(mats)
        draw[f"{tag}_iv"], draw[f"{tag}_ivp"] = iv, ivp
        draw[f"{tag}_mapsize"] = np.array(sc.map_size)
    np.savez_compressed(HERE / "llvmpipe_draw.npz", **draw)

    # ---- lighting
    light = {}
    cases = []
    sc = scenes.config_small()
    spec = scenes.config_small(seed=12)
    spec.materials[1].specular, spec.materials[1].shininess = 0.7, 3
    spec.materials[2].specular, spec.materials[2].shininess, spec.materials[2].reflect_type = 1.0, 100, 1
    for tag, scn, n_prev, nds, maxd, tm in (("first", sc, 0, 1, 1000, 123.456), ("rand1", sc, 1, 1, 1000, 200.25),
                                            ("rand5", sc, 5, 2, 4, 321.5), ("spec_first", spec, 0, 1, 1000, 123.456),
                                            ("spec_rand", spec, 3, 1, 1000, 250.75)):
        m, c, v, mats = build_state(scn)
        mm = m.copy()
        mm[(mm[:, 0] & 3) == 2, 0] |= 4
        cc = c.copy()
        cc[(mm[:, 0] & 3) == 2, 3] = n_prev
        req = o.visible_requests(mm, cc)
        s = ref.scene(scn.map_size, mm, cc, v, mats)
        vo, mo, co, dv, _ = ref.lighting(s, req, time=tm, num_diffuse=nds, max_diffuse=maxd, sun_dir=SUN,
                                         cam_pos=scn.cam_pos)
        # lit voxels: every stored voxel of every requested chunk
        idx = []
        for t in np.unique(req >> 4):
            n = int(sum(bin(int(w)).count("1") for w in cc[t, 7:23]))
            idx.extend(range(int(m[t, 2]), int(m[t, 2]) + n))
        idx = np.array(idx, np.int64)
        light[f"{tag}_map"], light[f"{tag}_chunks"], light[f"{tag}_voxels"] = mm, cc, v
        light[f"{tag}_mats"] = mats_u32(mats)
        light[f"{tag}_req"] = req
        light[f"{tag}_idx"] = idx
        light[f"{tag}_codes"] = vo[idx]
        light[f"{tag}_prequant"] = dv[idx]
        light[f"{tag}_flags_after"] = mo[:, 0].copy()
        light[f"{tag}_params"] = np.array([tm, nds, maxd, scn.cam_pos[0], scn.cam_pos[1], scn.cam_pos[2]], np.float64)
        seeds = frame_seeds(tm, nds, 5)
        light[f"{tag}_rand_seeds"] = seeds
        light[f"{tag}_rand_vals"] = ref.rand(seeds)
        light[f"{tag}_mapsize"] = np.array(scn.map_size)
        cases.append(tag)
    light["sun"] = SUN
    light["points"] = POINTS
    np.savez_compressed(HERE / "llvmpipe_light.npz", **light)
    (HERE / "llvmpipe_meta.json").write_text(json.dumps({"renderer": out["renderer"], "lighting_cases": cases,
                                                         "shaders": "assets/shaders/voxelDraw.comp, voxelLighting.comp, "
                                                                    "voxelShared.comp"}, indent=1) + "\n")
    print("wrote", HERE / "llvmpipe_draw.npz", HERE / "llvmpipe_light.npz", out["renderer"])


if __name__ == "__main__":
    main()

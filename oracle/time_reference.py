"""Times the reference path on the CPU: the reference's unmodified GLSL shaders on
Mesa llvmpipe (LP_NUM_THREADS threads), on the 128^3 and 512x256x512 scenes.

Ray and update counts come from the FP32 oracle on the same frame (llvmpipe
cannot count). Every time is the median of 10 frames after a warm-up frame
(shader JIT). The algorithm is the shaders' own, at their defaults: 15
specular rays, one jittered shadow ray per diffuse sample, <= 5 diffuse
bounces, 1 sample. Prints one JSON object. Needs the reference tree (shaders are
read from it at run time) and Mesa; run where both exist:
    LP_NUM_THREADS=8 python oracle/time_reference.py
"""
from __future__ import annotations

import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
from oracle.ref_llvmpipe import Llvmpipe  # noqa: E402
from doonengine_amd import scenes  # noqa: E402
from helpers import build_state, camera  # noqa: E402

SUN = np.array([-1.0, 1.0, -1.0], np.float32) * (np.float32(1.0) / np.sqrt(np.float32(3.0)))


def run(ref, o, sc, width, height, max_requests, frames=10, bounces=5):
    m, c, v, mats = build_state(sc)
    _, _, iv, ivp = camera(sc, width, height)
    s = ref.scene(sc.map_size, m, c, v, mats)
    ref.draw(s, iv, ivp, width, height)                          # warm-up (shader JIT)
    dts = [ref.draw(s, iv, ivp, width, height)[3] for _ in range(frames)]
    draw_s = float(np.median(dts))
    flags = m[:, 0].copy()
    os_ = o.scene(sc.map_size, m, c, v, mats)
    h16 = (height // 16) * 16                                    # the reference dispatches h/16 groups (voxel.c:879)
    _, _, dc = o.draw(os_, iv, ivp, width, h16, flags_out=flags)
    mm = m.copy()
    mm[:, 0] = flags
    req = o.visible_requests(mm, c)
    sub = req[:max_requests]
    pts = oracle.golden_points(15).astype(np.float32)
    s2 = ref.scene(sc.map_size, mm, c, v, mats)
    ref.lighting(s2, sub[: min(len(sub), 64)], time=1.0, sun_dir=SUN, diffuse_bounces=bounces, cam_pos=sc.cam_pos)
    lts = []
    for f in range(frames):
        cc = c.copy()
        cc[(mm[:, 0] & 3) == 2, 3] = 1 + f                      # past the first sample: random directions
        s3 = ref.scene(sc.map_size, mm, cc, v, mats)
        lts.append(ref.lighting(s3, sub, time=123.456 + f / 60, sun_dir=SUN, diffuse_bounces=bounces,
                                cam_pos=sc.cam_pos)[4])
    light_s = float(np.median(lts))
    cc = c.copy()
    cc[(mm[:, 0] & 3) == 2, 3] = 1
    _, _, _, _, voxof, lc = o.lighting(o.scene(sc.map_size, mm, cc, v, mats), sub, time=123.456, sun_dir=SUN,
                                       diffuse_bounces=bounces, spec_points=pts, cam_pos=sc.cam_pos)
    lit = int((voxof != 0xFFFFFFFF).sum())
    return {
        "draw": {"pixels": width * h16, "rays": dc["calls"], "seconds": draw_s, "rays_per_s": dc["calls"] / draw_s},
        "lighting": {"requests": int(len(sub)), "requests_in_frame": int(len(req)), "voxels_lit": lit,
                     "rays": lc["calls"], "seconds": light_s, "rays_per_s": lc["calls"] / light_s,
                     "updates_per_s": lit / light_s},
    }


def main():
    threads = int(os.environ.get("LP_NUM_THREADS", os.cpu_count() or 1))
    ref = Llvmpipe(instrument=False)
    o = oracle.load("f32")
    cpu = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")), "")
    out = {"renderer": ref.renderer, "threads": threads, "host_cpus": os.cpu_count(), "cpu": cpu, "frames": 10,
           "algorithm": "reference shaders as written: 15 specular rays, 1 jittered shadow ray per sample, "
                        "<= 5 diffuse bounces, 1 sample"}
    t0 = time.time()
    out["config1"] = run(ref, o, scenes.config1(reference_like=True), 256, 256, max_requests=1 << 30)
    out["config2_prefix4096"] = run(ref, o, scenes.config2(), 1920, 1080, max_requests=4096)
    out["wall_s"] = time.time() - t0
    print(json.dumps(out))


if __name__ == "__main__":
    main()

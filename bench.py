#!/usr/bin/env python3
"""Benchmark of the MI355X voxel pipeline on the full-frame workload.

One step = one frame of the engine loop (src/main.c:503-505): DN_draw of a
1920x1080 primary-ray cast, DN_sync_gpu (visible-chunk pass), DN_update_lighting
over the visible chunks. Workload: 64x32x64 chunks (512x256x512 voxels), fBm
terrain plus 2000 seeded boxes (~30% occupancy, ~64% empty chunks), 32 fixed
direct rays per diffuse sample, <= 4 diffuse bounces, FP32. The lighting load is
the configured "up to 32768 visible chunks": besides the chunks the draw marks
visible, DNX_mark_visible marks every exposed chunk (one holding a voxel with an
empty face neighbour; 16201 in this scene) visible each frame, a full relight of
everything that can be seen from somewhere.

Prints ONE JSON line. `value` is DDA rays/s (step_map calls of draw + lighting
per second); beside it:
- voxel-lighting updates/s, per-frame ray and step counts, per-kernel times;
- `roofline`: the lighting pass's algorithmic bytes / HIP-event duration against
  8 TB/s, and, from the committed rocprofv3 summary when it was taken of the
  same frame sources, its measured HBM traffic and fraction and the VALU
  counters that bound it;
- `cpu_baseline`: the FP32 oracle on a bounded sample of the frame, on this host's
  threads (at most 16);
- `reference_settings`: the same frame lit with the reference's own lighting
  algorithm (one jittered shadow ray per sample, 15 specular rays, <= 5
  bounces), the like-for-like figure for the llvmpipe baseline;
- `edits_summary` / `batch_summary`: 60 frames of the edit-streaming workload
  and 3 frames of the 384-volume batch (`--workload edits|batch` for the full runs).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
VISIBLE_CAP = 32768
PROFILE = ROOT / "profiles" / "frame_kernels_512x256x512_1920x1080.md"


FRAME_SOURCES = ("dn_device.hpp", "dn_internal.h", "dn_light.hpp", "engine.cpp", "kernels.hip")


def source_stamp() -> str:
    """First 16 hex digits of the SHA-256 of the sources the frame's kernels
    and their launches are built from (doonengine_amd/csrc), in name order:
    which sources a profile was taken of."""
    h = hashlib.sha256()
    for name in FRAME_SOURCES:
        f = ROOT / "doonengine_amd" / "csrc" / name
        h.update(name.encode())
        h.update(f.read_bytes())
    return h.hexdigest()[:16]


def profile_values() -> dict:
    """The `key: value` lines of the committed rocprofv3 summary of this
    workload (HBM bytes from FETCH_SIZE/WRITE_SIZE scaled by a plain copy,
    VALU counters), or {} when the summary was taken of other sources than
    the ones being benched."""
    vals = {}
    try:
        for line in PROFILE.read_text().splitlines():
            if ": " in line and not line.startswith(("|", "#", " ", "`")):
                k, v = line.split(": ", 1)
                vals[k.strip()] = v.strip()
    except OSError:
        return {}
    return vals if vals.get("source_sha16") == source_stamp() else {}


def fnum(vals: dict, key: str):
    try:
        return float(vals[key])
    except (KeyError, ValueError):
        return None


def algorithmic_bytes(c: dict) -> float:
    """Bytes per the traversal model: a chunk-grid trip reads a 12-byte tile,
    a voxel trip a 4-byte mask word, a hit 16 bytes of voxel plus 20 of rank
    data; each lit voxel reads 16, writes 12, plus its share of the chunk
    header (96/32) and request word (4/32)."""
    return (12.0 * c["mapSteps"] + 4.0 * c["voxelSteps"] + 36.0 * c["hits"]
            + c.get("litVoxels", 0) * (16.0 + 12.0 + 96.0 / 32.0 + 4.0 / 32.0))


def cpu_baseline(vol, sc, budget_s: float, threads: int) -> dict:
    """Times the FP32 oracle on `threads` host threads (request ranges and draw
    rows split over them) on a bounded prefix of the frame's lighting requests
    and a band of draw rows, on the exported device state."""
    import numpy as np

    import oracle
    from doonengine_amd.volume import draw_matrices, mat_to_np

    o = oracle.load("f32")
    out = __import__("doonengine_amd._hip", fromlist=["DeviceBuffer"]).DeviceBuffer(sc.width * sc.height * 16)
    view, proj = vol.view_projection(sc.height / sc.width)
    vol.draw(out, sc.width, sc.height, view, proj)
    vol.mark_visible(VISIBLE_CAP)
    vol.sync()
    req = vol.fetch_requests()
    m2, c2, v2 = vol.export_state()
    s = o.scene(sc.map_size, m2, c2, v2, vol.materials.copy(),
                sun=tuple(getattr(vol.v.sunStrength, k) for k in "xyz"))
    sd = np.array([vol.v.sunDir.x, vol.v.sunDir.y, vol.v.sunDir.z], np.float32)
    sd = sd * (np.float32(1.0) / np.sqrt(np.float32(sd[0] * sd[0] + sd[1] * sd[1]) + np.float32(sd[2] * sd[2])))
    kw = dict(time=float(vol.v.lastTime), diffuse_bounces=int(vol.v.diffuseBounceLimit), sun_dir=sd,
              spec_points=vol.ray_points(0), direct_points=vol.ray_points(1) if vol.v.numDirectRays else None,
              cam_pos=tuple(getattr(vol.v.camPos, k) for k in "xyz"))
    # Lighting: successive slices of the request list (doubling) until ~70% of
    # the budget of oracle compute time is spent; only the oracle call is timed.
    n, lt, lrays, lit, size = 0, 0.0, 0, 0, 16
    while n < len(req) and lt < 0.7 * budget_s:
        take = min(size, len(req) - n)
        _, _, _, _, voxof, lc = o.lighting(s, req[n:n + take], threads=threads, **kw)
        lt += lc["seconds"]
        lrays += lc["calls"]
        lit += int((voxof != 0xFFFFFFFF).sum())
        n += take
        size = min(size * 2, 4096)
    # Draw: the first rows of the frame (a row band costs about the same per ray).
    view, proj = vol.view_projection(sc.height / sc.width)
    iv, ivp = draw_matrices(view, proj)
    _, _, dc = o.draw(s, mat_to_np(iv), mat_to_np(ivp), sc.width, 8, threads=threads)
    rows = int(min(sc.height, max(8, 8 * 0.3 * budget_s / max(dc["seconds"], 1e-3))))
    if rows > 8:
        _, _, dc = o.draw(s, mat_to_np(iv), mat_to_np(ivp), sc.width, rows, threads=threads)
    dtime = dc["seconds"]
    rays = lrays + dc["calls"]
    return {"value": rays / (lt + dtime), "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"FP32 oracle, {threads} host thread(s): lighting of the first {n} of {len(req)} requests "
                      f"({lit} voxels, {lrays} rays, {lt:.1f} s) + draw of {rows} rows x {sc.width} "
                      f"({dc['calls']} rays, {dtime:.1f} s)",
            "lighting_updates_per_s": lit / max(lt, 1e-9)}


def bench_batch(args) -> int:
    print(json.dumps(run_batch(args.scenes, args.steps, args.warmup)), flush=True)
    return 0


def run_batch(n_scenes: int, steps: int, warmup: int) -> dict:
    """Batched uneven workload: `n_scenes` seeded volumes (chunk grids U[4,48]
    per axis, occupancy U[5%,60%], 2-16 materials, 10-40000 visible chunks),
    lit per frame either by one batched persistent graph launch (autotuned
    work split) or by one DN_update_lighting per volume. Both paths give
    bit-identical state (tests/test_gpu_batch.py, and every volume against the
    oracle in tests/test_gpu_configs.py); here they are timed on alternating
    frames of the same volumes."""
    from doonengine_amd import _hip, scenes
    from doonengine_amd.volume import KIND_LIGHTING, Batch, Volume

    t_setup = time.perf_counter()
    vols, vis = [], []
    for i in range(n_scenes):
        sc = scenes.batch_scene(i)
        v = Volume(sc.map_size, min_chunks=256)
        sc.apply(v)
        v.request_all()
        v.sync()
        vols.append(v)
        vis.append(sc.visible)
        if (i + 1) % 32 == 0:
            print(f"[bench] {i + 1} volumes built ({time.perf_counter() - t_setup:.0f} s)", file=sys.stderr, flush=True)
    _hip.synchronize()
    batch = Batch(vols)
    frame = [0]

    def frame_sync():
        for v, n in zip(vols, vis):
            v.mark_visible(n)
            v.sync()
        _hip.synchronize()
        frame[0] += 1
        return 123.456 + frame[0] / 60.0

    t = frame_sync()
    best, tune_ms = batch.autotune(1, 1000, t, repeats=2)
    t_tuned = time.perf_counter()
    print(f"[bench] autotuned {best} ({tune_ms:.2f} ms compute-only)", file=sys.stderr, flush=True)
    for _ in range(warmup):
        batch.update_lighting(1, 1000, t)
        t = frame_sync()
    for v in vols:
        v.counters(KIND_LIGHTING, reset=True)
    batch.timing(reset=True)
    tb = ts = 0.0
    lit_b = lit_s = 0
    for k in range(2 * steps):
        t0 = time.perf_counter()
        if k % 2 == 0:
            batch.update_lighting(1, 1000, t)
        else:
            for v in vols:
                v.update_lighting(1, 1000, t)
        _hip.synchronize()
        dt = time.perf_counter() - t0
        lit = sum(v.counters(KIND_LIGHTING, reset=True)["litVoxels"] for v in vols)
        if k % 2 == 0:
            tb, lit_b = tb + dt, lit_b + lit
        else:
            ts, lit_s = ts + dt, lit_s + lit
        t = frame_sync()
    ev_ms, launches = batch.timing(reset=True)
    K = steps
    result = {
        "metric": "voxel-lighting updates/s, batched uneven volumes (one persistent graph launch per frame)",
        "value": lit_b / tb, "unit": "updates/s", "n_gpus": 1, "steps": K, "warmup": warmup,
        "ms_per_step": 1e3 * tb / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n_scenes} seeded volumes, grids U[4,48]^3 chunks, occupancy U[5%,60%], "
                               "2-16 materials, 10-40000 visible chunks, 16 direct/specular rays, 1 diffuse "
                               "sample (<= 4 bounces)", "volumes": n_scenes, "parallelism": "single GPU"},
        "batched": {"ms_per_frame": 1e3 * tb / K, "graph_ms_per_frame": ev_ms / max(launches, 1),
                    "voxels_lit_per_frame": lit_b / K, "autotuned": best, "autotune_compute_ms": tune_ms},
        "per_volume_launches": {"ms_per_frame": 1e3 * ts / K, "voxels_lit_per_frame": lit_s / K,
                                "updates_per_s": lit_s / ts},
        "speedup": (ts / max(lit_s, 1)) / (tb / max(lit_b, 1)),
        "setup_s": t_tuned - t_setup,
    }
    batch.close()
    for v in vols:
        v.close()
    return result


def catmull_rom(points, u: float):
    """Closed Catmull-Rom spline through `points` at parameter u in [0, 1)."""
    import numpy as np
    n = len(points)
    x = (u % 1.0) * n
    i = int(x)
    t = x - i
    p0, p1, p2, p3 = (np.asarray(points[(i + k) % n], np.float64) for k in (-1, 0, 1, 2))
    pos = 0.5 * (2 * p1 + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t + (-p0 + 3 * p1 - 3 * p2 + p3) * t ** 3)
    tan = 0.5 * ((-p0 + p2) + 2 * (2 * p0 - 5 * p1 + 4 * p2 - p3) * t + 3 * (-p0 + 3 * p1 - 3 * p2 + p3) * t * t)
    return pos, tan


def bench_edits(args) -> int:
    print(json.dumps(run_edits(args.frames or 600, args.warmup, args.edit_api)), flush=True)
    return 0


def run_edits(frames: int, warmup: int, edit_api: str) -> dict:
    """Dynamic edit streaming on the 16^3-chunk fBm map: each frame 256 seeded
    sphere carve/fill edits (radius U[3,24], centre uniform over the map,
    50/50, seed 4 + frame), then the reference's frame order: DN_draw (256x256,
    camera on a seeded closed Catmull-Rom spline), DN_sync_gpu (re-uploads only
    the dirty chunks, whose lighting history restarts), DN_update_lighting.
    Frame time is host wall clock from the first edit to the end of the
    lighting pass."""
    import numpy as np

    from doonengine_amd import scenes
    from doonengine_amd._hip import DeviceBuffer
    from doonengine_amd.scenes import uniform01
    from doonengine_amd.volume import KIND_LIGHTING, Volume

    sc = scenes.config1()
    vol = Volume(sc.map_size, min_chunks=4096)
    sc.apply(vol)
    vol.request_all()
    vol.sync()
    vol.synchronize()
    W, H = sc.width, sc.height
    out = DeviceBuffer(W * H * 16)
    ctrl = [(1.0 + 14.0 * float(uniform01(k, 0, seed=4)), 7.0 + 6.0 * float(uniform01(k, 1, seed=4)),
             1.0 + 14.0 * float(uniform01(k, 2, seed=4))) for k in range(8)]
    # the edit stream is input data: drawn for every frame before timing starts
    frame_edits = []
    for f in range(frames + warmup):
        fe = scenes.sphere_edits(sc, f)
        frame_edits.append(Volume.sphere_edits(fe) if edit_api == "batch" else fe)
    times, edit_ms, up_bytes, up_chunks, lit = [], [], [], [], []
    vol.counters(KIND_LIGHTING, reset=True)
    vol.lib.DNX_get_upload_stats(vol.ptr, None, None, None, True)
    b, c_, v_ = (ctypes_u64() for _ in range(3))
    for f in range(frames + warmup):
        pos, tan = catmull_rom(ctrl, f / max(frames, 1))
        vol.set_vec("camPos", tuple(float(x) for x in pos))
        yaw = float(np.degrees(np.arctan2(tan[0], tan[2])))
        vol.set_vec("camOrient", (25.0, yaw, 0.0))
        t0 = time.perf_counter()
        if edit_api == "batch":
            vol.shape_spheres(frame_edits[f])
        else:
            for mat, alb, ctr, r, flip in frame_edits[f]:
                vol.shape_sphere(mat, alb, ctr, r, flip_normals=flip)
        t1 = time.perf_counter()
        view, proj = vol.view_projection(H / W)
        vol.draw(out, W, H, view, proj)
        vol.sync()
        vol.update_lighting(1, 1000, 123.456 + f / 60.0)
        vol.synchronize()
        t2 = time.perf_counter()
        vol.lib.DNX_get_upload_stats(vol.ptr, b, c_, v_, True)
        litf = vol.counters(KIND_LIGHTING, reset=True)["litVoxels"]
        if f >= warmup:
            times.append(t2 - t0)
            edit_ms.append(1e3 * (t1 - t0))
            up_bytes.append(b.value)
            up_chunks.append(c_.value)
            lit.append(litf)
        if (f + 1) % 100 == 0:
            print(f"[bench] {f + 1} frames", file=sys.stderr, flush=True)
    t = np.array(times)
    result = {
        "metric": "dynamic edit streaming: voxel-lighting updates/s and frame time p50/p99",
        "value": float(np.sum(lit) / np.sum(t)), "unit": "updates/s", "n_gpus": 1, "steps": frames,
        "warmup": warmup, "ms_per_step": float(1e3 * t.mean()), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "16x16x16 chunks (128^3 voxels) fBm ~35%, 256 sphere edits/frame (radius 3-24, "
                               "carve/fill 50/50), 256x256 draw, 16 direct/specular rays, 1 diffuse sample",
                   "frames": frames, "edit_api": "DNX_shape_spheres (device, one call per frame)"
                   if edit_api == "batch" else "DN_shape_sphere (host, one call per edit)",
                   "parallelism": "single GPU"},
        "frame_ms_p50": float(1e3 * np.percentile(t, 50)), "frame_ms_p99": float(1e3 * np.percentile(t, 99)),
        "edit_ms_p50": float(np.percentile(edit_ms, 50)),
        "upload_bytes_per_frame": float(np.mean(up_bytes)), "uploaded_chunks_per_frame": float(np.mean(up_chunks)),
        "dirty_fraction": float(np.mean(up_chunks) / (sc.map_size[0] * sc.map_size[1] * sc.map_size[2])),
        "voxels_lit_per_frame": float(np.mean(lit)),
    }
    vol.close()
    return result


def bench_cache(args) -> int:
    """Learned radiance cache vs plain accumulation on the 16^3-chunk fBm map
    (emission 1, so irradiance stays in the stored [0, 1] range), every exposed
    chunk lit each frame. Reference: per voxel, the FP64 mean of `--ref-samples`
    independent samples (per-sample random numbers, exact ambient + direct
    term), so that its own noise stays well below 1%. Each frame: the plain
    lighting pass (the engine's running mean, with the shader's frame-uniform
    random numbers) and one cache training step on 2^16 path-traced samples;
    both estimates are scored by relative RMSE (sqrt(mean squared error) / mean
    reference, over voxels and RGB)."""
    import numpy as np

    from doonengine_amd import scenes
    from doonengine_amd._hip import DeviceBuffer
    from doonengine_amd.volume import RadianceCache, Volume
    from oracle.radiance_cache import relative_rmse

    sc = scenes.config1(emission=1.0)
    vol = Volume(sc.map_size, min_chunks=4096)
    sc.apply(vol)
    vol.request_all()
    vol.sync()                      # loads every chunk
    vol.mark_visible(1 << 20)
    vol.sync()                      # the lighting request list of every exposed chunk
    vol.synchronize()
    n = vol.fetch_requests().size
    slots = n * 32
    rc = RadianceCache(vol, samplesPerFrame=args.cache_samples, seed=1, lr=args.cache_lr, emaDecay=args.cache_ema,
                       nMax=args.cache_nmax, log2Table=args.cache_table, lrDecaySteps=args.cache_decay_steps,
                       lrDecayPower=args.cache_decay_power, emaGrowth=args.cache_ema_growth)
    R, F = args.ref_samples, args.frames or 4096
    acc = DeviceBuffer(slots * 32)
    acc.zero()
    t0 = time.perf_counter()
    rc.reference(acc, n, R, seed=7)
    vol.synchronize()
    ref = acc.download(np.float64, (slots, 4))
    print(f"[bench] reference: {R} samples per voxel in {time.perf_counter() - t0:.1f} s, {n} requests",
          file=sys.stderr, flush=True)
    valid = ref[:, 3] > 0
    refm = np.clip(ref[valid, :3] / ref[valid, 3:4], 0.0, 1.0)

    dbg = DeviceBuffer(slots * 24)
    idx = DeviceBuffer(slots * 4)
    pred = DeviceBuffer(slots * 16)
    vol.lib.DNX_set_lighting_debug_output(vol.ptr, dbg.ptr, idx.ptr, n)
    # A second, cold accumulator started at frame F0: what an edit or a newly
    # visible region gets (its lighting history restarts), against the warm cache.
    F0 = F // 2
    cold = Volume(sc.map_size, min_chunks=4096)
    sc.apply(cold)
    cold.request_all()
    cold.sync()
    dbg2, idx2 = DeviceBuffer(slots * 24), DeviceBuffer(slots * 4)
    cold.lib.DNX_set_lighting_debug_output(cold.ptr, dbg2.ptr, idx2.ptr, n)
    curve_p, curve_c, curve_cold = [], [], []
    t_train = t_pred = 0.0
    for f in range(F):
        t = 123.456 + f / 60.0
        vol.mark_visible(1 << 20)
        vol.sync()
        vol.update_lighting(1, 1000, t)
        if f >= F0:
            cold.mark_visible(1 << 20)
            cold.sync()
            cold.update_lighting(1, 1000, t)
            cold.synchronize()
            curve_cold.append(relative_rmse(np.clip(dbg2.download(np.float32, (slots, 6))[:, :3][valid], 0, 1), refm))
        vol.synchronize()
        a = time.perf_counter()
        rc.train_frame(t, args.cache_steps)
        vol.synchronize()
        b = time.perf_counter()
        rc.predict_requests(pred, n)
        vol.synchronize()
        c = time.perf_counter()
        t_train += b - a
        t_pred += c - b
        plain = dbg.download(np.float32, (slots, 6))[:n * 32, :3][valid]
        cache = np.clip(pred.download(np.float32, (slots, 4))[:n * 32, :3][valid], 0, 1).astype(np.float64)
        curve_p.append(relative_rmse(np.clip(plain, 0, 1), refm))
        curve_c.append(relative_rmse(cache, refm))
        if (f + 1) % 32 == 0 or f < 8:
            print(f"[bench] frame {f + 1}: rel RMSE plain {curve_p[-1]:.4f} cache {curve_c[-1]:.4f}", file=sys.stderr,
                  flush=True)
    rc.close()
    vol.lib.DNX_set_lighting_debug_output(vol.ptr, None, None, 0)
    vol.close()
    cold.lib.DNX_set_lighting_debug_output(cold.ptr, None, None, 0)
    cold.close()

    def first_below(curve, thr=0.01):
        for i, e in enumerate(curve):
            if e <= thr:
                return i + 1
        return None

    # frames the cold accumulator needs to match the warm cache's error at F0
    warm = curve_c[F0 - 1] if F0 > 0 else None
    catch_up = first_below(curve_cold, warm) if warm is not None else None

    fp, fc = first_below(curve_p), first_below(curve_c)
    marks = [m for m in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096) if m <= F]
    result = {
        "metric": "frames to 1% relative RMSE: radiance cache vs plain accumulation",
        "value": fc, "unit": "frames", "n_gpus": 1, "steps": F, "warmup": 0,
        "ms_per_step": 1e3 * (t_train + t_pred) / F, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (MLP operands), f32 accumulate", "data": "synthetic",
        "config": {"workload": "16x16x16 chunks (128^3 voxels) fBm ~35%, every exposed chunk lit per frame, "
                               f"16 direct rays, 1 diffuse sample; cache: 16-level hash grid (2^{args.cache_table}), MLP 3x64, "
                               f"Adam lr {args.cache_lr}"
                               + (f" / (1 + step / {args.cache_decay_steps:g})^{args.cache_decay_power:g}"
                                  if args.cache_decay_steps > 0 else "")
                               + f", {args.cache_steps} step(s) per frame, each on {args.cache_samples} fresh "
                               f"samples, predictions from the parameter average (decay {args.cache_ema}"
                               + (f", window growing as step / {args.cache_ema_growth:g}" if args.cache_ema_growth > 0
                                  else "") + "); "
                               f"reference: {R} independent samples per voxel, FP64",
                   "voxels": int(valid.sum()), "parallelism": "single GPU"},
        "frames_to_1pct": {"cache": fc, "plain": fp},
        "frames_to": {f"{100 * thr:g}%": {"cache": first_below(curve_c, thr), "plain": first_below(curve_p, thr)}
                      for thr in (0.10, 0.05, 0.03, 0.02, 0.015, 0.01)},
        "rel_rmse_at": {str(m): {"plain": curve_p[m - 1], "cache": curve_c[m - 1]} for m in marks},
        "after_history_reset": {
            "reset_at_frame": F0, "warm_cache_rel_rmse": warm,
            "cold_plain_rel_rmse_after": {str(k): curve_cold[k - 1] for k in (1, 2, 4, 8, 16, 32, 64, 128)
                                          if k <= len(curve_cold)},
            "frames_for_cold_plain_to_match_warm_cache": catch_up},
        "train_ms_per_frame": 1e3 * t_train / F, "predict_ms_per_frame": 1e3 * t_pred / F,
    }
    print(json.dumps(result), flush=True)
    return 0


def ctypes_u64():
    import ctypes
    return ctypes.c_uint64()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("frame", "batch", "edits", "cache"), default="frame",
                    help="frame: the full-frame pipeline (default); batch: the batched uneven-volume lighting; "
                         "edits: dynamic edit streaming; cache: radiance cache vs plain accumulation")
    ap.add_argument("--scenes", type=int, default=384, help="volumes of the batch workload")
    ap.add_argument("--frames", type=int, default=None,
                    help="frames of the edits (default 600) and cache (default 4096) workloads")
    ap.add_argument("--edit-api", choices=["batch", "sphere"], default="batch",
                    help="edits workload: one DNX_shape_spheres call per frame (device) or DN_shape_sphere per edit (host)")
    ap.add_argument("--ref-samples", type=int, default=262144,
                    help="independent samples per voxel of the cache workload's FP64 reference")
    ap.add_argument("--cache-samples", type=int, default=1 << 17, help="cache training samples per step")
    ap.add_argument("--cache-steps", type=int, default=8,
                    help="cache Adam steps per frame, each on --cache-samples fresh samples")
    ap.add_argument("--cache-lr", type=float, default=1e-2, help="cache Adam learning rate")
    ap.add_argument("--cache-ema", type=float, default=0.9994, help="decay of the cache's parameter average")
    ap.add_argument("--cache-nmax", type=int, default=256, help="finest hash-grid resolution over the map")
    ap.add_argument("--cache-table", type=int, default=19, help="log2 of the hash-table entries per level")
    ap.add_argument("--cache-decay-steps", type=float, default=256.0,
                    help="cache learning-rate decay: lr / (1 + step / D)^P (0: constant)")
    ap.add_argument("--cache-decay-power", type=float, default=1.0, help="P of --cache-decay-steps")
    ap.add_argument("--cache-ema-growth", type=float, default=8.0,
                    help="cache parameter average: decay min(ema, 1 - G / step) while the window grows (0: off)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads of the CPU baseline (the oracle's request ranges and draw rows)")
    ap.add_argument("--ref-frames", type=int, default=5,
                    help="frame workload: frames timed with the reference's own lighting settings (0: skip)")
    ap.add_argument("--no-summaries", dest="summaries", action="store_false",
                    help="frame workload: skip the short edit-streaming and batched-volume summaries")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="after the timed frames, copy BYTES device-to-device once (a plain copy kernel of known "
                         "size, to scale profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "edits":
        return bench_edits(args)
    if args.workload == "cache":
        return bench_cache(args)

    import numpy as np

    from doonengine_amd import scenes
    from doonengine_amd._hip import DeviceBuffer
    from doonengine_amd.volume import KIND_DRAW, KIND_LIGHTING, Volume

    t_setup = time.perf_counter()
    sc = scenes.config2()
    vol = Volume(sc.map_size, min_chunks=65536)
    sc.apply(vol)
    vol.request_all()
    vol.sync()
    vol.synchronize()
    print(f"[bench] scene ready in {time.perf_counter() - t_setup:.1f} s: occupancy {sc.occupancy:.3f}, "
          f"pool {int(vol.v.voxelCap)} voxels", file=sys.stderr, flush=True)
    W, H = sc.width, sc.height
    out = DeviceBuffer(W * H * 16)
    view, proj = vol.view_projection(H / W)

    frame = [0]

    def step():
        vol.draw(out, W, H, view, proj)
        vol.mark_visible(VISIBLE_CAP)
        vol.sync()
        vol.update_lighting(1, 1000, 123.456 + frame[0] / 60.0)
        frame[0] += 1

    for _ in range(args.warmup):
        step()
    vol.synchronize()
    print(f"[bench] warmup done ({args.warmup} frames)", file=sys.stderr, flush=True)
    vol.counters(KIND_DRAW, reset=True)
    vol.counters(KIND_LIGHTING, reset=True)
    vol.enable_timing(True)
    vol.timing(reset=True)
    vol.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    vol.synchronize()
    elapsed = time.perf_counter() - t0
    vol.enable_timing(False)
    tm = vol.timing(reset=True)
    dc = vol.counters(KIND_DRAW)
    lc = vol.counters(KIND_LIGHTING)
    K = args.steps
    rays = dc["rays"] + lc["rays"]
    lit_per_frame = lc["litVoxels"] / K
    light_ms = tm["lightingMs"] / max(tm["lightingLaunches"], 1)
    light_bytes = algorithmic_bytes(lc) / K
    achieved = light_bytes / (light_ms * 1e-3) / 1e9
    print(f"[bench] {K} frames in {elapsed:.3f} s; lighting {light_ms:.3f} ms/launch, draw "
          f"{tm['drawMs'] / max(tm['drawLaunches'], 1):.3f} ms/launch, {lit_per_frame:.0f} voxels lit/frame",
          file=sys.stderr, flush=True)

    if args.copy_calibration > 0:
        from doonengine_amd import _hip
        a, b = DeviceBuffer(args.copy_calibration), DeviceBuffer(args.copy_calibration)
        a.zero()
        _hip.check(_hip.hip().hipMemcpy(b.ptr, a.ptr, args.copy_calibration, 3), "calibration copy")
        _hip.synchronize()
        print(f"[bench] calibration copy of {args.copy_calibration} bytes done", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        cpu = cpu_baseline(vol, sc, args.cpu_seconds, max(1, args.cpu_threads))

    # The same frame with the reference's own lighting algorithm
    # (voxelLighting.comp:207-288 at its defaults): one jittered shadow ray per
    # diffuse sample instead of the fixed direct rays, the 15 spherePoints for
    # specular rays, up to 5 diffuse bounces.
    refset = None
    if args.ref_frames > 0:
        vol.v.numDirectRays = 0
        vol.v.numSpecularRays = 15
        vol.v.diffuseBounceLimit = 5
        for _ in range(2):
            step()
        vol.synchronize()
        vol.counters(KIND_DRAW, reset=True)
        vol.counters(KIND_LIGHTING, reset=True)
        vol.enable_timing(True)
        vol.timing(reset=True)
        t1 = time.perf_counter()
        for _ in range(args.ref_frames):
            step()
        vol.synchronize()
        e1 = time.perf_counter() - t1
        vol.enable_timing(False)
        tm1 = vol.timing(reset=True)
        lc1 = vol.counters(KIND_LIGHTING)
        lms1 = tm1["lightingMs"] / max(tm1["lightingLaunches"], 1)
        refset = {"settings": "numDirectRays 0 (one jittered shadow ray per diffuse sample), 15 spherePoints "
                              "specular rays, diffuseBounceLimit 5, 1 diffuse sample",
                  "frames": args.ref_frames, "ms_per_frame": 1e3 * e1 / args.ref_frames, "lighting_ms": lms1,
                  "voxel_lighting_updates_per_s": lc1["litVoxels"] / (lms1 * 1e-3 * args.ref_frames),
                  "lighting_rays_per_s": lc1["rays"] / (lms1 * 1e-3 * args.ref_frames),
                  "lighting_rays_per_update": lc1["rays"] / max(lc1["litVoxels"], 1)}
        print(f"[bench] reference settings: lighting {lms1:.3f} ms/launch", file=sys.stderr, flush=True)

    result = {
        "metric": "DDA rays/s & voxel-lighting updates/s; cache vs plain frames-to-1%-RMSE",
        "value": rays / elapsed,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "full frame: 64x32x64 chunks (512x256x512 voxels), fBm terrain + 2000 boxes, "
                        "1920x1080 primary DDA draw, lighting over <= 32768 visible chunks, 32 direct/specular "
                        "rays, 1 diffuse sample (<= 4 bounces), FP32",
            "map_chunks": list(sc.map_size),
            "image": [W, H],
            "visible_chunks_cap": VISIBLE_CAP,
            "parallelism": "single GPU",
        },
        "voxel_lighting_updates_per_s": lc["litVoxels"] / elapsed,
        "draw_rays_per_s": dc["rays"] / elapsed,
        "lighting_rays_per_s": lc["rays"] / elapsed,
        "per_frame": {"draw_rays": dc["rays"] / K, "lighting_rays": lc["rays"] / K, "voxels_lit": lit_per_frame,
                      "lighting_map_steps": lc["mapSteps"] / K, "lighting_voxel_steps": lc["voxelSteps"] / K,
                      "lighting_hits": lc["hits"] / K, "draw_map_steps": dc["mapSteps"] / K,
                      "draw_voxel_steps": dc["voxelSteps"] / K},
        "kernel_ms": {"lighting": light_ms, "draw": tm["drawMs"] / max(tm["drawLaunches"], 1),
                      "visible": tm["visibleMs"] / max(tm["visibleLaunches"], 1),
                      "epilogue": tm["epilogueMs"] / max(tm["epilogueLaunches"], 1)},
        "roofline": roofline(light_bytes, light_ms, achieved),
        "cpu_baseline": cpu,
        "reference_settings": refset,
    }
    vol.close()
    # short summaries of the edit-streaming and batched-volume workloads
    if args.summaries:
        ed = run_edits(60, 5, "batch")
        result["edits_summary"] = {k: ed[k] for k in ("value", "unit", "frame_ms_p50", "frame_ms_p99", "edit_ms_p50",
                                                       "upload_bytes_per_frame", "voxels_lit_per_frame")}
        result["edits_summary"]["workload"] = ed["config"]["workload"] + ", 60 frames"
        print(f"[bench] edits summary: frame p50 {ed['frame_ms_p50']:.2f} ms", file=sys.stderr, flush=True)
        bt = run_batch(384, 3, 1)
        result["batch_summary"] = {"value": bt["value"], "unit": bt["unit"], "workload": bt["config"]["workload"],
                                   "batched_ms_per_frame": bt["batched"]["ms_per_frame"],
                                   "per_volume_ms_per_frame": bt["per_volume_launches"]["ms_per_frame"],
                                   "speedup": bt["speedup"], "voxels_lit_per_frame": bt["batched"]["voxels_lit_per_frame"],
                                   "frames": 3}
        print(f"[bench] batch summary: {bt['batched']['ms_per_frame']:.1f} ms/frame", file=sys.stderr, flush=True)
    print(json.dumps(result), flush=True)
    return 0


def roofline(light_bytes: float, light_ms: float, achieved: float) -> dict:
    """The lighting pass against the HBM roofline on the traversal model's
    algorithmic bytes, with what the committed profile of the same kernel
    sources measured beside it: the real HBM traffic (so the fraction of HBM bandwidth actually
    used) and the VALU counters that show what bounds the kernels."""
    pv = profile_values()
    traffic = fnum(pv, "traffic_bytes_per_launch lighting_pass")
    out = {"kernel": "lighting pass (k_direct + k_lighting)", "bound": "hbm", "achieved": achieved,
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
           "algorithmic_bytes_per_launch": light_bytes,
           "profile_source_sha16": pv.get("source_sha16"), "source_sha16": source_stamp()}
    if traffic is not None:
        measured = traffic / (light_ms * 1e-3) / 1e9
        out["measured_hbm_gbs"] = measured
        out["measured_hbm_frac"] = measured / HBM_PEAK_GBS
    valu = {}
    for k in ("k_direct", "k_lighting"):
        busy, lane = fnum(pv, f"valu_busy {k}"), fnum(pv, f"lane_use {k}")
        if busy is not None:
            valu[k] = {"valu_busy": busy, "lane_use": lane, "valu_insts": fnum(pv, f"valu_insts {k}")}
    if valu:
        valu["lane_insts_per_dda_trip"] = fnum(pv, "valu_lane_insts_per_step lighting_pass")
        out["limiter"] = ("VALU issue: the kernels' VALU pipes are busy most of the time while HBM traffic "
                          "stays far below peak (the scene stays in L2); lanes idle on rays that end early")
        out["valu"] = valu
    return out


if __name__ == "__main__":
    sys.exit(main())

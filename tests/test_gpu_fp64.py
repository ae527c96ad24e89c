"""GPU lighting against the FP64 oracle (same algorithm in double precision).

Deterministic terms (ambient + fixed direct rays with the diffuse path off,
and specular with the fixed sphere points) agree within 1e-4 relative on
every voxel.
Diffuse path tracing is compared over N frames: the GPU's mean of its
per-frame samples against the FP64 oracle's, per voxel, inside a 3-sigma
Monte-Carlo band, sigma being the voxel's FP64 per-frame sample standard
deviation / sqrt(N). Both draw the same frame-uniform random numbers, so the
band is loose; the measured errors are printed.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from doonengine_amd import scenes
from doonengine_amd._hip import DeviceBuffer
from test_gpu_parity import load_scene, oracle_scene, sun_dir_normalized

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4
FLOOR = 1e-2      # relative errors are taken against max(|fp64 value|, FLOOR)


def light_kw(vol, num_diffuse=1):
    return dict(time=float(vol.v.lastTime), num_diffuse=num_diffuse, max_diffuse=1000,
                diffuse_bounces=int(vol.v.diffuseBounceLimit), spec_bounces=int(vol.v.specBounceLimit),
                sun_dir=sun_dir_normalized(vol),
                ambient=tuple(getattr(vol.v.ambientLightStrength, k) for k in "xyz"),
                softness=float(vol.v.shadowSoftness), cam_pos=tuple(getattr(vol.v.camPos, k) for k in "xyz"),
                spec_points=vol.ray_points(0), direct_points=vol.ray_points(1) if vol.v.numDirectRays else None)


def specular_scene():
    sc = scenes.config_small(seed=12)
    sc.materials[1].specular = 0.7
    sc.materials[1].shininess = 3
    sc.materials[2].specular = 1.0
    sc.materials[2].shininess = 100
    sc.materials[2].reflect_type = 1
    return sc


def test_deterministic_terms_within_1e4(gpu):
    sc = specular_scene()
    o64 = oracle.load("f64")
    vol = load_scene(sc)
    vol.v.numDirectRays = 8
    vol.v.numSpecularRays = 15
    vol.v.diffuseBounceLimit = 0
    vol.mark_visible(1 << 20)
    vol.sync()
    req = vol.fetch_requests()
    R = len(req)
    dbg, dbi = DeviceBuffer(R * 32 * 24), DeviceBuffer(R * 32 * 4)
    dbg.zero()                       # slots without a voxel are never written
    vol.lib.DNX_set_lighting_debug_output(vol.ptr, dbg.ptr, dbi.ptr, R)
    s, _ = oracle_scene(o64, vol, sc)
    vol.update_lighting(1, 1000, 300.0)
    vol.synchronize()
    vol.lib.DNX_set_lighting_debug_output(vol.ptr, None, None, 0)
    g = dbg.download(np.float32, (R * 32, 6)).astype(np.float64)
    _, _, _, preq, voxof, _ = o64.lighting(s, req, **light_kw(vol))
    valid = voxof != 0xFFFFFFFF
    err = np.abs(g[valid] - preq[valid]) / np.maximum(np.abs(preq[valid]), FLOOR)
    per_voxel = err.max(axis=1)
    frac_ok = float((per_voxel <= REL_TOL).mean())
    spec_vox = (np.abs(preq[valid][:, 3:]).max(axis=1) > 0).sum()
    print(f"[fp64] deterministic: {valid.sum()} voxels ({spec_vox} with specular), rel err median "
          f"{np.median(per_voxel):.2e} max {per_voxel.max():.2e}, within 1e-4: {frac_ok:.5f}", flush=True)
    assert spec_vox > 50
    # every voxel (measured max 3.4e-7): no DDA decision of these rays differs between FP32 and FP64
    assert frac_ok == 1.0
    vol.close()


def test_diffuse_within_mc_band(gpu):
    sc = scenes.config_small()
    o64 = oracle.load("f64")
    vol = load_scene(sc)
    vol.v.numDirectRays = 8
    vol.v.diffuseBounceLimit = 4
    N = 32
    vol.mark_visible(1 << 20)
    vol.sync()
    req = vol.fetch_requests()
    R = len(req)
    fresh = DeviceBuffer(R * 32 * 16)
    vol.lib.DNX_set_lighting_sample_output(vol.ptr, fresh.ptr, R)
    m, c, v = vol.export_state()
    gs = np.zeros((N, R * 32, 3))
    os_ = np.zeros((N, R * 32, 3))
    ofresh = np.zeros((R * 32, 3), np.float64)
    o64.set_fresh_output(ofresh)
    try:
        state = (m, c, v)
        for f in range(N):
            t = 400.0 + f / 60.0
            vol.mark_visible(1 << 20)
            vol.sync()
            assert np.array_equal(vol.fetch_requests(), req)
            vol.update_lighting(1, 1000, t)
            vol.synchronize()
            gs[f] = fresh.download(np.float32, (R * 32, 4))[:, :3]
            s, _ = oracle_scene(o64, vol, sc, state)
            vo, mo, co, _, voxof, _ = o64.lighting(s, req, **light_kw(vol))
            os_[f] = ofresh
            state = (mo, co, vo)
    finally:
        o64.set_fresh_output(None)
        vol.lib.DNX_set_lighting_sample_output(vol.ptr, None, 0)
    valid = voxof != 0xFFFFFFFF
    gm, om = gs[:, valid].mean(axis=0), os_[:, valid].mean(axis=0)
    # Paired: both sides draw the same random directions, so frame by frame the
    # samples differ only where FP32 and FP64 take different DDA decisions on a
    # grazing ray (then that frame's path differs entirely). The GPU estimate is
    # unbiased against FP64 when the mean difference stays within 3 standard
    # errors of the per-frame differences (plus FP32 rounding, 1e-4 relative).
    d = gs[:, valid] - os_[:, valid]
    se = d.std(axis=0, ddof=1) / np.sqrt(N)
    band = 3.0 * se + REL_TOL * np.maximum(np.abs(om), FLOOR)
    inside = np.abs(gm - om) <= band
    frac = float(inside.all(axis=1).mean())
    diverged = float((np.abs(d) > REL_TOL * np.maximum(np.abs(os_[:, valid]), FLOOR)).any(axis=2).mean())
    rel = np.abs(gm - om) / np.maximum(np.abs(om), FLOOR)
    sig = os_[:, valid].std(axis=0, ddof=1) / np.sqrt(N)
    print(f"[fp64] diffuse over {N} frames: {valid.sum()} voxels; voxel-frames whose path differs from FP64 "
          f"{diverged:.4f}; N-frame mean vs FP64: mean abs err {np.abs(gm - om).mean():.2e}, median rel "
          f"{np.median(rel):.2e}, max rel {rel.max():.2e}; MC standard error of the estimate (mean) {sig.mean():.2e}; "
          f"within 3 paired standard errors: {frac:.5f}", flush=True)
    assert frac >= 0.997
    vol.close()

"""GPU parity: the HIP kernels through the C-ABI against the FP32 oracle.

Each check exports the device state (in the reference's buffer layouts) right
before a pass, runs the oracle on exactly that state, and compares:
bit-exact for hit voxel index, face, step counts, request lists, stored
lighting codes and pre-quantisation FP32 values; draw colours within 2e-6
relative (the only non-bit-exact operation there is powf).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from doonengine_amd import scenes
from doonengine_amd._hip import DeviceBuffer
from doonengine_amd.volume import PIXELHIT_DTYPE, Volume, mat_to_np
from helpers import build_state, camera, materials_array

pytestmark = pytest.mark.gpu


def load_scene(sc, min_chunks=64):
    vol = Volume(sc.map_size, min_chunks=min_chunks)
    sc.apply(vol)
    vol.request_all()
    vol.sync()
    vol.synchronize()
    return vol


def oracle_scene(o, vol, sc, state=None):
    m, c, v = state if state is not None else vol.export_state()
    mats = vol.materials.copy()
    s = o.scene(sc.map_size, m, c, v, mats, sun=tuple(getattr(vol.v.sunStrength, k) for k in "xyz"),
                sky_bot=tuple(getattr(vol.v.skyGradientBot, k) for k in "xyz"),
                sky_top=tuple(getattr(vol.v.skyGradientTop, k) for k in "xyz"))
    if vol.v.useCubemap and getattr(vol, "cube_faces", None) is not None:
        o.set_cubemap(s, vol.cube_faces)
    return s, (m, c, v)


def draw_gpu(vol, sc, w=None, h=None):
    w = w or sc.width
    h = h or sc.height
    view, proj = vol.view_projection(h / w, 0.1, 100.0)
    out = DeviceBuffer(w * h * 16)
    hits = DeviceBuffer(w * h * PIXELHIT_DTYPE.itemsize)
    vol.lib.DNX_set_draw_hit_output(vol.ptr, hits.ptr, w * h)
    vol.draw(out, w, h, view, proj)
    vol.synchronize()
    vol.lib.DNX_set_draw_hit_output(vol.ptr, None, 0)
    rgba = out.download(np.float32, (h, w, 4))
    px = hits.download(PIXELHIT_DTYPE, (h, w))
    return rgba, px, view, proj


def sun_dir_normalized(vol):
    d = np.array([vol.v.sunDir.x, vol.v.sunDir.y, vol.v.sunDir.z], np.float32)
    n = np.float32(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1]) + np.float32(d[2] * d[2])))
    return d * (np.float32(1.0) / n) if n != 0 else np.zeros(3, np.float32)


@pytest.fixture(scope="module")
def small_scene():
    return scenes.config_small()


def test_upload_matches_oracle_conversion(gpu, small_scene):
    sc = small_scene
    vol = load_scene(sc)
    m, c, v = vol.export_state()
    rm, rc, rv, _ = build_state(sc, materials_array(sc))
    loaded = (m[:, 0] & 3) == 2
    assert np.array_equal(loaded, (rm[:, 0] & 3) == 2)
    assert np.array_equal(c[loaded], rc[loaded])          # pos, samples 0, partialCounts, bitMask
    for t in np.nonzero(loaded)[0]:
        n = int(sum(bin(int(w)).count("1") for w in c[t, 7:23]))
        assert np.array_equal(v[m[t, 2]:m[t, 2] + n], rv[rm[t, 2]:rm[t, 2] + n])
    vol.close()


def test_draw_parity(gpu, small_scene):
    sc = small_scene
    o = oracle.load("f32")
    vol = load_scene(sc)
    s, (m, c, v) = oracle_scene(o, vol, sc)
    rgba, px, view, proj = draw_gpu(vol, sc)
    from doonengine_amd.volume import draw_matrices
    iv, ivp = draw_matrices(view, proj)
    flags = m[:, 0].copy()
    orgba, opx, ocnt = o.draw(s, mat_to_np(iv), mat_to_np(ivp), sc.width, sc.height, flags_out=flags)
    assert (px["voxelIndex"] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(px["voxelIndex"], opx["voxelIndex"])
    assert np.array_equal(px["mapSteps"], opx["mapSteps"])
    assert np.array_equal(px["voxelSteps"], opx["voxelSteps"])
    assert np.array_equal(px["face"].view(np.uint32) & 0x7FFFFFFF, opx["face"].view(np.uint32) & 0x7FFFFFFF)
    assert np.array_equal(rgba[..., 3], orgba[..., 3])    # depth: sqrt and arithmetic only
    np.testing.assert_allclose(rgba[..., :3], orgba[..., :3], rtol=2e-6, atol=1e-7)
    m2, _, _ = vol.export_state()
    assert np.array_equal(m2[:, 0], flags)                # visible bits
    cnt = vol.counters(0, reset=True)
    assert cnt["rays"] == ocnt["calls"] and cnt["mapSteps"] == ocnt["sMap"] and cnt["voxelSteps"] == ocnt["sVox"]
    vol.close()


def test_rays_parity(gpu, small_scene):
    sc = small_scene
    o = oracle.load("f32")
    vol = load_scene(sc)
    s, _ = oracle_scene(o, vol, sc)
    rng = np.random.default_rng(5)
    n = 20000
    X, Y, Z = sc.map_size
    org = rng.uniform(-0.5, 1.0, (n, 3)) * np.array([X, Y, Z]) + 0.25
    d = rng.normal(size=(n, 3))
    # axis-parallel and diagonal directions (ties in the DDA)
    d[: n // 10] = np.eye(3)[rng.integers(0, 3, n // 10)] * rng.choice([-1, 1], (n // 10, 1))
    d[n // 10: n // 5] = rng.choice([-1.0, 1.0], (n // 10, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = (d + 1e-4).astype(np.float32)
    org = org.astype(np.float32)
    # integer-aligned origins (start exactly on cell boundaries)
    org[-1000:] = np.floor(org[-1000:])
    for ignore in (False, True):
        g = vol.cast_rays(org, d, ignore_first=ignore)
        r, _ = o.cast_rays(s, org, d, ignore_first=ignore)
        for f in ("hit", "voxelIndex", "mapSteps", "voxelSteps"):
            assert np.array_equal(g[f], r[f]), f
        assert np.array_equal(g["pos"].view(np.uint32), r["pos"].view(np.uint32))
        assert np.array_equal(g["normal"].view(np.uint32) & 0x7FFFFFFF, r["normal"].view(np.uint32) & 0x7FFFFFFF)
        assert g["hit"].mean() > 0.1
    vol.close()


def oracle_lighting(o, vol, s, req, num_diffuse=1, max_diffuse=1000, threads=1):
    """The oracle's lighting dispatch on scene `s` with the volume's own lighting
    parameters (the time it latched, bounce limits, sun, ambient, softness,
    camera, specular and direct ray tables)."""
    direct = vol.ray_points(1) if vol.v.numDirectRays else None
    return o.lighting(
        s, req, time=float(vol.v.lastTime), num_diffuse=num_diffuse, max_diffuse=max_diffuse,
        diffuse_bounces=int(vol.v.diffuseBounceLimit), spec_bounces=int(vol.v.specBounceLimit),
        sun_dir=sun_dir_normalized(vol),
        ambient=tuple(getattr(vol.v.ambientLightStrength, k) for k in "xyz"), softness=float(vol.v.shadowSoftness),
        cam_pos=tuple(getattr(vol.v.camPos, k) for k in "xyz"), spec_points=vol.ray_points(0), direct_points=direct,
        threads=threads)


def run_frame_and_compare(vol, sc, o, frame, time, w, h, num_diffuse=1, max_diffuse=1000, split=1, threads=1,
                          mark_visible=0, full_draw=False, states=None):
    """draw -> sync(READ_WRITE) -> lighting on the GPU, each pass checked against
    the oracle (`threads` host threads for the oracle's passes). The draw's and
    the lighting rays' requests of unloaded chunks (flags = 3) are part of the
    compared flags. mark_visible > 0 also marks up to that many exposed chunks
    visible after the draw, as the full-frame workload does. full_draw compares
    the whole draw record (face, trip counts, depth, colour), not only the hit
    voxels. `states`, a dict, receives the exported state after the draw (m1,
    c1) and after the sync (m2, c2)."""
    vol.counters(0, reset=True)                            # counts of earlier, unchecked passes
    vol.counters(1, reset=True)
    s, (m0, c0, v0) = oracle_scene(o, vol, sc)
    rgba, px, view, proj = draw_gpu(vol, sc, w, h)
    from doonengine_amd.volume import draw_matrices
    iv, ivp = draw_matrices(view, proj)
    flags = m0[:, 0].copy()
    s.flagsOut = flags.ctypes.data                         # ray requests (flags = 3) land in the same words
    orgba, opx, dcnt = o.draw(s, mat_to_np(iv), mat_to_np(ivp), w, h, flags_out=flags, threads=threads)
    s.flagsOut = None
    assert np.array_equal(px["voxelIndex"], opx["voxelIndex"])
    if full_draw:
        assert np.array_equal(px["mapSteps"], opx["mapSteps"])
        assert np.array_equal(px["voxelSteps"], opx["voxelSteps"])
        assert np.array_equal(px["face"].view(np.uint32) & 0x7FFFFFFF, opx["face"].view(np.uint32) & 0x7FFFFFFF)
        assert np.array_equal(rgba[..., 3], orgba[..., 3])
        np.testing.assert_allclose(rgba[..., :3], orgba[..., :3], rtol=2e-6, atol=1e-7)
        dc = vol.counters(0, reset=True)
        assert dc["rays"] == dcnt["calls"] and dc["mapSteps"] == dcnt["sMap"] and dc["voxelSteps"] == dcnt["sVox"]
    # visible-chunk pass
    m1, c1, v1 = vol.export_state()
    assert np.array_equal(m1[:, 0], flags)
    if mark_visible:
        vol.mark_visible(mark_visible)
        m1, c1, v1 = vol.export_state()
    vol.sync(op=2, split=split)
    req = vol.fetch_requests()
    ref_req = o.visible_requests(m1, c1, split=split, frame=int(vol.v.frameNum))
    assert np.array_equal(req, ref_req)
    m2, c2, v2 = vol.export_state()
    # lastUsed++ on every tile (voxel.c:753), except chunks (re)loaded by this sync, which restart at 0
    # (voxel.c:1526); every chunk that became loaded is one of them
    aged = m2[:, 1] == m1[:, 1] + 1
    restarted = ~aged
    assert np.all(m2[restarted, 1] == 0) and np.all((m2[restarted, 0] & 3) == 2)
    newly = ((m2[:, 0] & 3) == 2) & ((m1[:, 0] & 3) != 2)
    assert np.all(restarted[newly])
    if states is not None:
        states.update(m1=m1, c1=c1, m2=m2, c2=c2)
    # lighting pass
    R = max(len(req), 1)
    dbg = DeviceBuffer(R * 32 * 6 * 4)
    dbi = DeviceBuffer(R * 32 * 4)
    vol.lib.DNX_set_lighting_debug_output(vol.ptr, dbg.ptr, dbi.ptr, R)
    s2, _ = oracle_scene(o, vol, sc, (m2, c2, v2))
    vol.update_lighting(num_diffuse, max_diffuse, time)
    vol.synchronize()
    vol.lib.DNX_set_lighting_debug_output(vol.ptr, None, None, 0)
    g_val = dbg.download(np.float32, (R * 32, 6))[: len(req) * 32]
    g_idx = dbi.download(np.uint32, (R * 32,))[: len(req) * 32]
    vo, mo, co, preq, voxof, ocnt = oracle_lighting(o, vol, s2, req, num_diffuse, max_diffuse, threads)
    m3, c3, v3 = vol.export_state()
    assert np.array_equal(g_idx, voxof)
    valid = voxof != 0xFFFFFFFF
    assert np.array_equal(g_val[valid].view(np.uint32), preq[valid].astype(np.float32).view(np.uint32))
    assert np.array_equal(v3, vo)
    assert np.array_equal(m3[:, 0], mo[:, 0])
    assert np.array_equal(c3[:, 3], co[:, 3])
    loaded = (m3[:, 0] & 3) == 2
    assert np.array_equal(m3[loaded, 1], mo[loaded, 1])    # lastUsed of loaded chunks
    cnt = vol.counters(1, reset=True)
    assert cnt["rays"] == ocnt["calls"] and cnt["mapSteps"] == ocnt["sMap"] and cnt["voxelSteps"] == ocnt["sVox"]
    assert cnt["litVoxels"] == int(valid.sum())
    return len(req), int(valid.sum())


def test_lighting_parity_frames(gpu, small_scene):
    sc = small_scene
    o = oracle.load("f32")
    vol = load_scene(sc)
    total = 0
    for f in range(4):
        nreq, nlit = run_frame_and_compare(vol, sc, o, f, 123.456 + f / 60.0, sc.width, sc.height)
        total += nlit
    assert total > 1000
    vol.close()


def test_lighting_parity_specular_and_direct(gpu):
    sc = scenes.config_small(seed=12)
    sc.materials[1].specular = 0.7
    sc.materials[1].shininess = 3
    sc.materials[2].specular = 1.0
    sc.materials[2].shininess = 100
    sc.materials[2].reflect_type = 1
    o = oracle.load("f32")
    vol = load_scene(sc)
    vol.v.numDirectRays = 8
    vol.v.numSpecularRays = 15
    for f in range(3):
        run_frame_and_compare(vol, sc, o, f, 200.0 + f / 60.0, sc.width, sc.height)
    vol.v.numSpecularRays = 16
    run_frame_and_compare(vol, sc, o, 3, 201.0, sc.width, sc.height)
    vol.close()


def test_lighting_split_and_samples(gpu, small_scene):
    sc = small_scene
    o = oracle.load("f32")
    vol = load_scene(sc)
    for f in range(4):
        run_frame_and_compare(vol, sc, o, f, 300.0 + f, sc.width, sc.height, num_diffuse=2, max_diffuse=2, split=3)
    vol.close()


def test_transparent_draw_parity(gpu):
    sc = scenes.config_small(seed=13)
    sc.materials[0].opacity = 0.4
    o = oracle.load("f32")
    vol = load_scene(sc)
    vol.materials[0]["refractIndex"] = 1.3
    s, (m, c, v) = oracle_scene(o, vol, sc)
    rgba, px, view, proj = draw_gpu(vol, sc)
    from doonengine_amd.volume import draw_matrices
    iv, ivp = draw_matrices(view, proj)
    orgba, opx, _ = o.draw(s, mat_to_np(iv), mat_to_np(ivp), sc.width, sc.height)
    assert np.array_equal(px["voxelIndex"], opx["voxelIndex"])
    assert np.array_equal(px["mapSteps"], opx["mapSteps"])
    assert np.array_equal(px["voxelSteps"], opx["voxelSteps"])
    np.testing.assert_allclose(rgba[..., :3], orgba[..., :3], rtol=2e-6, atol=1e-7)
    vol.close()


def test_lighting_parity_direct_ray_with_zero_component(gpu):
    """A direct-ray direction with an exactly-zero component (infinite delta on
    that axis, where the reference's 0 * delta is NaN) takes the reference walk;
    the other directions take the finite-delta walk. Both against the oracle.
    Sun along -x with softness 1e-4 turns the first direct point (0, 1, 0) into
    normalize(-1e-4, 1, 0) + 1e-4 = (0, 1.0001, 1e-4)."""
    sc = scenes.config_small(seed=14)
    o = oracle.load("f32")
    vol = load_scene(sc)
    vol.v.numDirectRays = 8
    vol.v.sunDir.x, vol.v.sunDir.y, vol.v.sunDir.z = -1.0, 0.0, 0.0
    vol.v.shadowSoftness = 1e-4
    q = vol.ray_points(1).astype(np.float32)
    f = np.float32
    v = f(-1.0) * f(1e-4) + q[:, 0], f(0.0) * f(1e-4) + q[:, 1], f(0.0) * f(1e-4) + q[:, 2]
    inv = f(1.0) / np.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
    dirs = np.stack([c * inv + f(1e-4) for c in v], axis=1)
    assert (dirs == 0).any(axis=1).sum() >= 1 and (dirs != 0).all(axis=1).sum() >= 4
    for fr in range(2):
        run_frame_and_compare(vol, sc, o, fr, 400.0 + fr / 60.0, sc.width, sc.height)
    vol.close()


def test_cubemap_sky_draw_and_lighting_parity(gpu):
    """Sky from a cubemap (DNX_set_sky_cubemap) in the draw's background and in
    the lighting pass's mirror reflections (reflectType 1 specular rays that
    miss), against the oracle's GL-rules sampler."""
    sc = scenes.config_small(seed=15)
    sc.materials[1].specular = 1.0
    sc.materials[1].shininess = 100
    sc.materials[1].reflect_type = 1
    o = oracle.load("f32")
    vol = load_scene(sc)
    rng = np.random.default_rng(15)
    vol.set_sky_cubemap(rng.integers(0, 256, (6, 16, 16, 3), dtype=np.uint8))
    vol.v.numSpecularRays = 15
    s, _ = oracle_scene(o, vol, sc)
    rgba, px, view, proj = draw_gpu(vol, sc)
    from doonengine_amd.volume import draw_matrices
    iv, ivp = draw_matrices(view, proj)
    orgba, opx, _ = o.draw(s, mat_to_np(iv), mat_to_np(ivp), sc.width, sc.height)
    sky = px["voxelIndex"] == 0xFFFFFFFF
    assert sky.mean() > 0.05
    assert np.array_equal(px["voxelIndex"], opx["voxelIndex"])
    np.testing.assert_allclose(rgba[..., :3], orgba[..., :3], rtol=2e-6, atol=1e-7)
    for f in range(2):
        run_frame_and_compare(vol, sc, o, f, 500.0 + f / 60.0, sc.width, sc.height)
    vol.close()


@pytest.mark.parametrize("n_direct", [1, 24, 64])
def test_lighting_parity_direct_ray_counts(gpu, n_direct):
    """Direct-ray counts that fill a wave unevenly (64 / N voxels x N directions
    per wave: 64 x 1, 2 x 24 with 16 idle lanes, 1 x 64)."""
    sc = scenes.config_small(seed=16)
    o = oracle.load("f32")
    vol = load_scene(sc)
    vol.v.numDirectRays = n_direct
    run_frame_and_compare(vol, sc, o, 0, 600.0, sc.width, sc.height)
    run_frame_and_compare(vol, sc, o, 1, 600.0 + 1 / 60.0, sc.width, sc.height)
    vol.close()


@pytest.mark.parametrize("size", [(97, 61), (8, 8), (1, 200), (333, 7)])
def test_draw_parity_odd_image_sizes(gpu, size):
    """The persistent draw takes 8 x 8 quads: images whose sides are not
    multiples of 8, a single quad, a one-pixel column."""
    sc = scenes.config_small(seed=17)
    o = oracle.load("f32")
    vol = load_scene(sc)
    w, h = size
    s, (m, c, v) = oracle_scene(o, vol, sc)
    rgba, px, view, proj = draw_gpu(vol, sc, w, h)
    from doonengine_amd.volume import draw_matrices
    iv, ivp = draw_matrices(view, proj)
    flags = m[:, 0].copy()
    orgba, opx, ocnt = o.draw(s, mat_to_np(iv), mat_to_np(ivp), w, h, flags_out=flags)
    for f in ("voxelIndex", "mapSteps", "voxelSteps"):
        assert np.array_equal(px[f], opx[f]), f
    np.testing.assert_allclose(rgba[..., :3], orgba[..., :3], rtol=2e-6, atol=1e-7)
    m2, _, _ = vol.export_state()
    assert np.array_equal(m2[:, 0], flags)
    cnt = vol.counters(0, reset=True)
    assert cnt["rays"] == ocnt["calls"] == w * h
    vol.close()


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5])
def test_draw_view_modes_parity(gpu, mode):
    """DN_draw's other view modes (voxelDraw.comp:32-61: albedo, diffuse light,
    specular light, voxel normal, face normal) on a lit scene with a glossy
    material, against the oracle's draw in the same mode."""
    sc = scenes.config_small(seed=14)
    sc.materials[1].specular = 0.5
    sc.materials[1].shininess = 8
    o = oracle.load("f32")
    vol = load_scene(sc)
    vol.v.numDirectRays = 4
    for f in range(2):                               # stored diffuse and specular light that is not zero
        rgba0, _, _, _ = draw_gpu(vol, sc)
        vol.mark_visible(1 << 20)
        vol.sync()
        vol.update_lighting(1, 1000, 5.0 + f)
    vol.synchronize()
    vol.v.camViewMode = mode
    s, (m, _, v) = oracle_scene(o, vol, sc)
    assert (v[:, 3] != 0).any()
    rgba, px, view, proj = draw_gpu(vol, sc)
    from doonengine_amd.volume import draw_matrices
    iv, ivp = draw_matrices(view, proj)
    orgba, opx, _ = o.draw(s, mat_to_np(iv), mat_to_np(ivp), sc.width, sc.height, view_mode=mode,
                           flags_out=m[:, 0].copy())
    assert (px["voxelIndex"] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(px["voxelIndex"], opx["voxelIndex"])
    assert np.array_equal(rgba[..., 3], orgba[..., 3])
    np.testing.assert_allclose(rgba[..., :3], orgba[..., :3], rtol=2e-6, atol=1e-7)
    assert not np.allclose(rgba[..., :3], rgba0[..., :3])      # the mode changes the image
    vol.close()

"""GPU parity at the full size of each workload, against the oracle.

- 16x16x16 chunks (128^3 voxels, fBm), 256x256 draw, 16 direct and specular
  rays, emission 4: frames 1-3 and frame 64 of the accumulation, every pass
  compared (draw record and colour, request list, stored lighting words,
  pre-quantisation values, flags, sample counts, ray counters).
- 64x32x64 chunks (512x256x512 voxels, fBm + 2000 boxes), 1920x1080 draw,
  lighting of every chunk the draw sees plus up to 32768 exposed chunks, 32
  direct and specular rays, <= 4 diffuse bounces: two whole frames (the first
  sample's deterministic directions, then random ones), compared the same way (the oracle's passes split over host threads, which gives the
  single-thread result exactly).
- Edit streaming on the 16^3-chunk map: two frames of 256 seeded sphere carves
  and fills of radius 3-24 through DNX_shape_spheres, against the oracle's
  sequential edit model voxel for voxel; then the sync's upload (chunk records,
  stored voxels) against the oracle's chunk conversion, the lighting history of
  the re-uploaded chunks restarted, and a lighting frame on the edited map.
- The batched workload at its full size: 384 seeded volumes (chunk grids
  U[4,48]^3, 10-40000 visible chunks each) lit by one batched graph launch per
  frame after autotuning, every volume's lit voxel words, flags, sample counts
  and ray counters against the oracle, on two frames (the first sample's
  deterministic directions, then the frame-uniform random ones).
- Radiance cache at its full size: a 2^19-entry table per level and a batch of
  2^16 samples, forward outputs and every gradient against the oracle.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from doonengine_amd import scenes
from doonengine_amd._hip import DeviceBuffer
from doonengine_amd.volume import KIND_LIGHTING, Batch, RadianceCache, Volume, rc_pack_weights, rc_unpack_weights
from oracle import radiance_cache as orc
from test_gpu_parity import oracle_lighting, oracle_scene, run_frame_and_compare

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def load(sc, min_chunks):
    vol = Volume(sc.map_size, min_chunks=min_chunks)
    sc.apply(vol)
    vol.request_all()
    vol.sync()
    vol.synchronize()
    return vol


@pytest.mark.timeout(600)
def test_c1_128cube_frames(gpu):
    sc = scenes.config1()
    assert sc.params["numDirectRays"] == 16 and sc.materials[-1].emission == 4.0
    o = oracle.load("f32")
    vol = load(sc, 4096)
    try:
        lit = 0
        for f in range(3):
            _, n = run_frame_and_compare(vol, sc, o, f, 123.456 + f / 60.0, sc.width, sc.height, threads=THREADS,
                                         full_draw=True)
            lit += n
        for f in range(3, 63):                       # accumulate to frame 64 without checking
            view, proj = vol.view_projection(sc.height / sc.width, 0.1, 100.0)
            out = DeviceBuffer(sc.width * sc.height * 16)
            vol.draw(out, sc.width, sc.height, view, proj)
            vol.sync()
            vol.update_lighting(1, 1000, 123.456 + f / 60.0)
        vol.synchronize()
        _, n = run_frame_and_compare(vol, sc, o, 63, 123.456 + 63 / 60.0, sc.width, sc.height, threads=THREADS,
                                     full_draw=True)
        _, c, _ = vol.export_state()
        assert c[:, 3].max() >= 64                    # the visible chunks hold 64 samples
        print(f"[c1] 4 checked frames, {lit + n} voxels lit", flush=True)
        assert lit > 30000
    finally:
        vol.close()


@pytest.fixture(scope="module")
def c2_volume(gpu):
    sc = scenes.config2()
    vol = load(sc, 65536)
    yield vol, sc
    vol.close()


@pytest.mark.timeout(900)
def test_c2_full_frame(c2_volume):
    vol, sc = c2_volume
    o = oracle.load("f32")
    assert (sc.width, sc.height) == (1920, 1080)
    # frame 1: every voxel's first sample (deterministic directions); frame 2: the frame-uniform random ones
    for f in range(2):
        nreq, nlit = run_frame_and_compare(vol, sc, o, f, 123.456 + f / 60.0, sc.width, sc.height, threads=THREADS,
                                           mark_visible=32768, full_draw=True)
        print(f"[c2] frame {f}: {nreq} requests, {nlit} voxels lit, all bit-exact", flush=True)
        assert nreq > 100000 and nlit > 3_000_000


def oracle_upload(o, grid, mats, t, map_size):
    X, Y, Z = map_size
    x, y, z = t % X, (t // X) % Y, t // (X * Y)
    sl = (slice(z * 8, z * 8 + 8), slice(y * 8, y * 8 + 8), slice(x * 8, x * 8 + 8))
    cv = np.stack([grid.normal[sl].transpose(2, 1, 0), grid.albedo[sl].transpose(2, 1, 0)], axis=-1)
    ch, vv = o.chunk_to_gpu(mats, cv)
    ch[0:3] = (x, y, z)
    return ch, vv


@pytest.mark.timeout(600)
def test_c3_edit_frames(gpu):
    sc = scenes.config1()
    o = oracle.load("f32")
    vol = load(sc, 4096)
    try:
        X, Y, Z = (8 * c for c in sc.map_size)
        T = int(np.prod(sc.map_size))
        grid = o.edit_grid(sc.map_size)
        grid.set_dense(sc.normal_words, sc.albedo_words)
        n0, a0 = vol.get_dense((X, Y, Z))
        np.testing.assert_array_equal(n0, grid.normal)
        vol.mark_visible(1 << 20)                     # light every exposed chunk once: a history to restart
        vol.sync()
        vol.update_lighting(1, 1000, 100.0)
        vol.synchronize()
        _, c_lit, _ = vol.export_state()
        assert (c_lit[:, 3] == 1).sum() > 1000
        for f in range(2):
            edits = scenes.sphere_edits(sc, f)
            assert len(edits) == 256 and min(e[3] for e in edits) >= 3.0 and max(e[3] for e in edits) <= 24.0
            vol.shape_spheres(edits)
            for mat, alb, c, r, flip in edits:
                grid.shape_sphere(mat, alb, c, r, flip_normals=flip)
            n1, a1 = vol.get_dense((X, Y, Z))
            np.testing.assert_array_equal(n1, grid.normal, err_msg=f"normal words after frame {f}")
            np.testing.assert_array_equal(a1, grid.albedo, err_msg=f"albedo words after frame {f}")
            flags = np.array([vol.chunk_exists((t % sc.map_size[0], (t // sc.map_size[0]) % sc.map_size[1],
                                                t // (sc.map_size[0] * sc.map_size[1]))) for t in range(T)])
            np.testing.assert_array_equal(flags, grid.flag.reshape(-1).astype(bool))
            m_before, _, _ = vol.export_state()
            vol.residency_stats(reset=True)
            vol.sync()
            vol.synchronize()
            m, c, v = vol.export_state()
            mats = vol.materials.copy()
            loaded = (m[:, 0] & 3) == 2
            present = grid.flag.reshape(-1) != 0
            # every CPU chunk is on the GPU, or unloaded (evicted as least recently used, or added and
            # waiting for a ray's request)
            assert np.all(loaded <= present)
            assert np.all((m[present & ~loaded, 0] & 3) == 1)
            assert np.all((m[~present, 0] & 3) == 0)
            # a loaded chunk that is still on the CPU stays loaded (re-uploaded when edited) unless its
            # node went to another chunk: only chunks no ray entered in the last two syncs
            # (voxel.c:1580-1604); without a pool limit the pool grows instead of deferring a load
            was = (m_before[:, 0] & 3) == 2
            lost = was & present & ~loaded
            stats = vol.residency_stats(reset=True)
            assert np.all(m_before[lost, 1] >= 2)
            assert stats["evictions"] == int(lost.sum()) and stats["deferrals"] == 0
            assert np.all(loaded <= was)                   # nothing requested a new chunk in between
            dirty = np.zeros(T, bool)
            for mat, alb, ctr, r, flip in edits:
                lo = np.maximum(np.floor((np.array(ctr) - r - 1) / 8).astype(int), 0)
                hi = np.minimum(np.floor((np.array(ctr) + r + 1) / 8).astype(int), np.array(sc.map_size) - 1)
                for zz in range(lo[2], hi[2] + 1):
                    for yy in range(lo[1], hi[1] + 1):
                        for xx in range(lo[0], hi[0] + 1):
                            dirty[xx + sc.map_size[0] * (yy + sc.map_size[1] * zz)] = True
            checked = fresh = 0
            for t in np.nonzero(loaded)[0]:
                ch, vv = oracle_upload(o, grid, mats, int(t), sc.map_size)
                assert np.array_equal(c[t, 4:23], ch[4:23]), f"chunk record of tile {t}"
                assert np.array_equal(c[t, 0:3], ch[0:3])
                n = len(vv)
                assert np.array_equal(v[m[t, 2]:m[t, 2] + n, 0:2], vv[:, 0:2]), f"voxels of tile {t}"
                if c[t, 3] == 0:                       # re-uploaded: lighting history restarted
                    assert np.all(v[m[t, 2]:m[t, 2] + n, 2:4] == 0)
                    fresh += 1
                checked += 1
            assert fresh > 100 and fresh <= int(dirty.sum())
            print(f"[c3] frame {f}: {checked} chunks checked, {fresh} re-uploaded", flush=True)
        nreq, nlit = run_frame_and_compare(vol, sc, o, 2, 200.0, sc.width, sc.height, threads=THREADS)
        assert nlit > 1000
    finally:
        vol.close()


@pytest.mark.timeout(900)
def test_c4_batch_384_volumes(gpu):
    o = oracle.load("f32")
    vols, scs = [], []
    batch = None
    try:
        for i in range(384):
            sc = scenes.batch_scene(i)
            v = Volume(sc.map_size, min_chunks=256)
            sc.apply(v)
            v.request_all()
            v.sync()
            vols.append(v)
            scs.append(sc)
            if (i + 1) % 64 == 0:
                print(f"[c4] {i + 1} volumes built", flush=True)
        for v in vols:
            v.synchronize()
        batch = Batch(vols)
        for f in range(2):
            t = 123.456 + f / 60.0
            for v, sc in zip(vols, scs):
                v.mark_visible(sc.visible)
                v.sync()
                v.counters(KIND_LIGHTING, reset=True)
            if f == 0:
                batch.autotune(1, 1000, t, repeats=1)
            before = []
            for v in vols:
                v.synchronize()
                before.append((v.export_state(), v.fetch_requests()))
            batch.update_lighting(1, 1000, t)
            lit = 0
            for i, (v, sc) in enumerate(zip(vols, scs)):
                v.synchronize()
                (m2, c2, v2), req = before[i]
                before[i] = None
                s2, _ = oracle_scene(o, v, sc, (m2, c2, v2))
                vo, mo, co, _, voxof, ocnt = oracle_lighting(o, v, s2, req, threads=THREADS)
                m3, c3, v3 = v.export_state()
                assert np.array_equal(v3, vo), f"frame {f}: voxel words of volume {i}"
                assert np.array_equal(m3[:, 0], mo[:, 0]), f"frame {f}: flags of volume {i}"
                assert np.array_equal(c3[:, 3], co[:, 3]), f"frame {f}: sample counts of volume {i}"
                cnt = v.counters(KIND_LIGHTING, reset=True)
                assert (cnt["rays"], cnt["mapSteps"], cnt["voxelSteps"]) == (ocnt["calls"], ocnt["sMap"], ocnt["sVox"])
                n = int((voxof != 0xFFFFFFFF).sum())
                assert cnt["litVoxels"] == n
                lit += n
            print(f"[c4] frame {f}: 384 volumes, {lit} voxels lit, all bit-exact", flush=True)
            assert lit > 50_000_000
    finally:
        if batch is not None:
            batch.close()
        for v in vols:
            v.close()


def make_samples(rng, n, extent=128.0):
    s = np.zeros((n, 16), np.float32)
    s[:, 0:3] = rng.uniform(0.5, extent - 0.5, (n, 3))
    nrm = rng.normal(size=(n, 3))
    s[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    s[:, 6:9] = rng.uniform(0.1, 0.9, (n, 3))
    s[:, 9:12] = rng.uniform(0.0, 1.0, (n, 3))
    s[:, 12] = 1.0
    return s


@pytest.mark.timeout(600)
def test_c5_full_size_gradients(gpu):
    sc = scenes.config1(emission=1.0)
    vol = load(sc, 4096)
    cfg = orc.Config(log2_table=19, extent=128.0)
    rng = np.random.default_rng(19)
    mats, table = orc.init_params(cfg, seed=9)
    table = rng.normal(0, 0.05, table.shape).astype(np.float32)
    B = 1 << 16
    smp = make_samples(rng, B)
    rc = RadianceCache(vol, log2Table=19, extent=128.0)
    try:
        rc.set_params(rc_pack_weights(mats), table.reshape(-1))
        buf = DeviceBuffer.from_array(smp)
        out = DeviceBuffer(B * 16)
        rc.predict_samples(buf, B, out)
        rc.train_samples(buf, B, step=False)
        vol.synchronize()
        y = out.download(np.float32, (B, 4))
        _, _, g, tg = rc.get_params()
        ref = orc.forward_backward(cfg, mats, table, smp[:, 0:3], smp[:, 3:6], smp[:, 6:9], smp[:, 9:12],
                                   bf16_ops=True)
        # bf16 operands: an FP32 sum taken in another order can round an activation to the neighbouring
        # bf16 value (2^-8 relative), so single outputs may differ by a few bf16 steps; the norms agree closely
        d = y[:, :3] - ref["y"]
        ey_max = float(np.abs(d).max() / np.abs(ref["y"]).max())
        ey = float(np.linalg.norm(d) / np.linalg.norm(ref["y"]))
        print(f"[c5] forward rel err {ey:.2e} (max {ey_max:.2e})", flush=True)
        ew = [float(np.linalg.norm(gm - rm) / np.linalg.norm(rm)) for gm, rm in zip(rc_unpack_weights(g), ref["dW"])]
        print(f"[c5] dW rel err {' '.join(f'{e:.2e}' for e in ew)}", flush=True)
        dt = tg.reshape(table.shape)
        et = float(np.linalg.norm(dt - ref["dtable"]) / np.linalg.norm(ref["dtable"]))
        touched = ref["dtable"] != 0
        print(f"[c5] dtable rel err {et:.2e}, {int(touched.sum())} entries touched", flush=True)
        assert ey < 2e-3 and ey_max < 4 * 2.0 ** -8
        assert max(ew) < 5e-3
        assert et < 5e-3
        assert np.all(dt[~touched.any(axis=-1)] == 0.0)
    finally:
        rc.close()
        vol.close()

"""Ray-requested streaming on the GPU, against the oracle (voxelShared.comp
:462-466, voxel.c:1491-1536).

A volume whose chunks are all on the CPU but none on the GPU (no
DNX_request_all_chunks): every frame's draw and lighting rays that enter an
unloaded chunk write its flag 3, exactly where the oracle's rays do; the next
DN_sync_gpu uploads exactly those chunks (with the records the oracle's chunk
conversion gives), and the frame after draws and lights them bit for bit like
the oracle. The voxel pool starts at 512 voxels, so it grows while chunks
stream in.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from doonengine_amd import scenes
from doonengine_amd.volume import Volume
from helpers import build_state, materials_array
from test_gpu_parity import run_frame_and_compare

pytestmark = pytest.mark.gpu


def test_ray_requested_streaming(gpu):
    sc = scenes.config_small(seed=21)
    o = oracle.load("f32")
    vol = Volume(sc.map_size, min_chunks=1)
    try:
        sc.apply(vol)
        assert vol.v.voxelCap == 512
        vol.sync()
        vol.synchronize()
        m, c, v = vol.export_state()
        rm, rc, rv, _ = build_state(sc, materials_array(sc))
        present = (rm[:, 0] & 3) == 2
        assert np.array_equal((m[:, 0] & 3) == 1, present)           # on the CPU, not on the GPU
        assert not ((m[:, 0] & 3) == 2).any()
        loaded_before = np.zeros(len(m), bool)
        requested_total = 0
        for f in range(6):
            st = {}
            run_frame_and_compare(vol, sc, o, f, 123.456 + f / 60.0, sc.width, sc.height, states=st)
            m1, m2, c2 = st["m1"], st["m2"], st["c2"]
            req = (m1[:, 0] & 3) == 3
            requested_total += int(req.sum())
            loaded = (m2[:, 0] & 3) == 2
            # the sync uploads exactly the requested chunks; loaded chunks stay
            assert np.array_equal(loaded, loaded_before | req), f"frame {f}"
            assert np.array_equal(c2[loaded, 4:23], rc[loaded, 4:23])
            assert np.array_equal(c2[loaded, 0:3], rc[loaded, 0:3])
            m3, _, _ = vol.export_state()
            loaded_before = (m3[:, 0] & 3) == 2
        assert requested_total > 5 and loaded_before.sum() > 5
        assert vol.v.voxelCap > 512                                   # the pool grew while streaming
        _, _, v = vol.export_state()
        for t in np.nonzero(loaded_before)[0]:
            n = int(sum(bin(int(w)).count("1") for w in rc[t, 7:23]))
            assert np.array_equal(v[m3[t, 2]:m3[t, 2] + n, 0:2], rv[rm[t, 2]:rm[t, 2] + n, 0:2])
    finally:
        vol.close()


def test_bounded_pool_evicts_least_recently_used(gpu):
    """A voxel pool bounded below what the map needs, with the camera turning
    around the map: chunks that no ray entered for two syncs are evicted to
    make room (voxel.c:1554-1611), loads that find no room wait for a later
    request, the pool never outgrows its bound, and every frame still draws
    and lights bit for bit like the oracle on the state the GPU holds."""
    sc = scenes.config_small(seed=22, chunks=(8, 3, 8), width=48, height=40)
    sc.cam_pos = (4.0, 2.6, 4.0)
    o = oracle.load("f32")
    rm, rc, _, _ = build_state(sc, materials_array(sc))
    present = (rm[:, 0] & 3) == 2
    need = 0
    for t in np.nonzero(present)[0]:
        n = int(sum(bin(int(w)).count("1") for w in rc[t, 7:23]))
        need += 16 << max(0, int(np.ceil(np.log2(max(n, 16) / 16))))
    limit = (need // 3 + 511) // 512 * 512
    vol = Volume(sc.map_size, min_chunks=1)
    try:
        sc.apply(vol)
        assert vol.set_pool_limit(limit)
        vol.sync()
        evicted_seen = deferred_seen = 0
        for f in range(16):
            vol.set_vec("camOrient", (35.0, 45.0 * f, 0.0))
            st = {}
            run_frame_and_compare(vol, sc, o, f, 123.456 + f / 60.0, sc.width, sc.height, states=st)
            m1, m2 = st["m1"], st["m2"]
            was = (m1[:, 0] & 3) == 2
            now = (m2[:, 0] & 3) == 2
            evicted = was & ~now
            assert np.all(m1[evicted, 1] >= 2), f"frame {f}: evicted a chunk used in the last two syncs"
            assert np.all((m2[evicted, 0] & 3) == 1)
            requested = (m1[:, 0] & 3) == 3
            deferred = requested & ~now
            assert np.all((m2[deferred, 0] & 3) == 1)
            assert np.all(now <= (was | requested))
            evicted_seen += int(evicted.sum())
            deferred_seen += int(deferred.sum())
            assert vol.v.voxelCap <= limit
            loaded = now
            c2 = st["c2"]
            assert np.array_equal(c2[loaded, 4:23], rc[loaded, 4:23])
        stats = vol.residency_stats()
        assert stats["evictions"] == evicted_seen and stats["deferrals"] == deferred_seen
        assert evicted_seen > 0
        assert vol.v.voxelCap > 512 and not vol.set_pool_limit(512)   # below the pool's current size
        assert vol.set_pool_limit(limit)
        print(f"[stream] pool {vol.v.voxelCap} of limit {limit} (map needs {need}), {evicted_seen} evictions, "
              f"{deferred_seen} deferred loads", flush=True)
    finally:
        vol.close()

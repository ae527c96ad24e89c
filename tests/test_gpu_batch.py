"""Batched lighting (DNX_batch_*) against per-volume DN_update_lighting.

Two identical sets of seeded uneven volumes are built; one set is lit volume by
volume, the other in one batched launch per frame. After every frame the
exported device state (map, chunk records with sample counts, voxel pool) and
the ray/step counters must be identical bit for bit, for several work-split
configurations and after an autotuning pass (which must leave state alone).
"""
from __future__ import annotations

import numpy as np
import pytest

from doonengine_amd import scenes
from doonengine_amd.volume import KIND_LIGHTING, Batch, Volume

pytestmark = pytest.mark.gpu

N_SCENES = 10
MAX_GRID = 10


def build_set():
    vols, vis = [], []
    for i in range(N_SCENES):
        sc = scenes.batch_scene(i, max_grid=MAX_GRID)
        v = Volume(sc.map_size, min_chunks=64)
        sc.apply(v)
        v.request_all()
        v.sync()
        vols.append(v)
        vis.append(sc.visible)
    for v in vols:
        v.synchronize()
    return vols, vis


def frame_sync(vols, vis):
    for v, n in zip(vols, vis):
        v.mark_visible(n)
        v.sync()


def assert_same(va, vb, tag, before=None):
    for i, (a, b) in enumerate(zip(va, vb)):
        a.synchronize()
        b.synchronize()
        for k, (name, x, y) in enumerate(zip(("map", "chunks", "voxels"), a.export_state(), b.export_state())):
            if not np.array_equal(x, y):
                rows = np.nonzero((x != y).reshape(len(x), -1).any(axis=1))[0]
                msg = f"{tag}: volume {i} {name} differs in {len(rows)} of {len(x)} rows, first {rows[:8]}"
                if before is not None:
                    pb = before[i][k]
                    msg += (f"; unchanged from before the pass: a {int((x[rows] == pb[rows]).all(axis=-1).sum())}, "
                            f"b {int((y[rows] == pb[rows]).all(axis=-1).sum())}")
                raise AssertionError(msg)
        ca, cb = a.counters(KIND_LIGHTING, reset=True), b.counters(KIND_LIGHTING, reset=True)
        assert ca == cb, f"{tag}: volume {i} counters {ca} != {cb}"


@pytest.mark.parametrize("configs", [
    [(8, 256, 0, 1)],
    [(2, 256, 0, 1)],
    [(8, 64, 0, 1)],
    [(8, 256, 0, 0)],
    [(2, 64, 0, 0), (32, 256, 1, 1), (4, 64, 2, 1)],
])
def test_batch_matches_sequential(gpu, configs):
    va, vis = build_set()
    vb, _ = build_set()
    batch = Batch(vb)
    try:
        lit_total = 0
        for f in range(3):
            t = 123.456 + f / 60.0
            frame_sync(va, vis)
            frame_sync(vb, vis)
            for v in vb:
                v.synchronize()
            before = [v.export_state() for v in vb]
            for v in va:
                v.update_lighting(1, 1000, t)
            batch.set_config(*configs[f % len(configs)])
            batch.update_lighting(1, 1000, t)
            lit_total += sum(v.counters(KIND_LIGHTING)["litVoxels"] for v in va)
            assert_same(va, vb, f"frame {f} config {configs[f % len(configs)]}", before)
        assert lit_total > 0
        print(f"[batch] {N_SCENES} volumes, 3 frames, {lit_total} voxels lit, identical", flush=True)
    finally:
        batch.close()
        for v in va + vb:
            v.close()


def test_autotune_leaves_state(gpu):
    va, vis = build_set()
    vb, _ = build_set()
    batch = Batch(vb)
    try:
        t = 200.0
        frame_sync(va, vis)
        frame_sync(vb, vis)
        best, ms = batch.autotune(1, 1000, t, repeats=1)
        assert ms > 0.0
        assert best["requestsPerUnit"] % 2 == 0 and best["blockThreads"] in (64, 256)
        assert batch.config == best
        for v in va:
            v.update_lighting(1, 1000, t)
        batch.update_lighting(1, 1000, t)
        assert_same(va, vb, f"after autotune {best}")
        ms_run, n = batch.timing(reset=True)
        assert n == 1 and ms_run > 0.0
    finally:
        batch.close()
        for v in va + vb:
            v.close()


def test_autotune_leaves_unloaded_chunks_alone(gpu):
    """Autotuning runs compute-only passes: with chunks present on the CPU but
    not loaded (a roof added after the first sync), whose tiles the lighting
    rays enter, the dry passes must neither request them (flags 3)
    nor stamp lastUsed; the real pass that follows does request them."""
    vols, vis = build_set()
    try:
        for v in vols:
            # a one-voxel roof over the whole map: the shadow rays that leave the
            # terrain upwards cross its (new, unloaded) chunks
            X, Y, Z = (8 * c for c in v.map_size)
            nw = np.full((Z, 1, X), (1 << 24) | (127 << 16) | (255 << 8) | 127, np.uint32)
            aw = np.full((Z, 1, X), 0xC8C8C800, np.uint32)
            v.set_dense(nw, aw, origin=(0, Y - 1, 0))
        frame_sync(vols, vis)
        for v in vols:
            v.synchronize()
        before = [v.export_state() for v in vols]
        assert any(((m[:, 0] & 3) == 1).any() for m, _, _ in before)       # unloaded chunks exist
        batch = Batch(vols)
        try:
            batch.autotune(1, 1000, 321.0, repeats=1)
            for v, (m0, c0, p0) in zip(vols, before):
                m1, c1, p1 = v.export_state()
                assert np.array_equal(m1, m0) and np.array_equal(c1, c0) and np.array_equal(p1, p0)
            batch.update_lighting(1, 1000, 321.0)
            requested = sum(int(((v.export_state()[0][:, 0] & 3) == 3).sum()) for v in vols)
            assert requested > 0
        finally:
            batch.close()
    finally:
        for v in vols:
            v.close()

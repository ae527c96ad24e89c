"""Edge cases on the GPU, against the oracle through the C-ABI.

- An empty map: every primary ray misses (sky only), the visible pass emits
  no request, and the lighting pass over no request changes nothing.
- A one-chunk map.
- Ragged request blocks: chunks whose stored-voxel counts are 1, 31, 32, 33
  and 296 (a solid chunk after per-chunk face culling, voxel.c:1393), so the
  last 32-voxel request block of a chunk is full, ragged or a single voxel.
- Lighting arguments outside the documented limits (more diffuse samples or
  bounces than the kernels hold, no specular rays): refused with an ERROR
  message, and the device state is left as it was.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from doonengine_amd import scenes
from doonengine_amd.scenes import EMPTY, Scene, pack_normal_word
from doonengine_amd.volume import Volume
from test_gpu_parity import run_frame_and_compare

pytestmark = pytest.mark.gpu


def empty_words(map_size):
    X, Y, Z = (8 * c for c in map_size)
    return np.full((Z, Y, X), np.uint32(EMPTY) << np.uint32(24), np.uint32), np.zeros((Z, Y, X), np.uint32)


def solid(nw, aw, z, y, x, material=0, normal=(0.0, 1.0, 0.0), albedo=(200, 150, 100)):
    n = np.array(normal, np.float64)
    nw[z, y, x] = pack_normal_word(np.uint32(material), n / np.linalg.norm(n))
    aw[z, y, x] = (albedo[0] << 24) | (albedo[1] << 16) | (albedo[2] << 8)


def load(sc):
    vol = Volume(sc.map_size, min_chunks=64)
    sc.apply(vol)
    vol.request_all()
    vol.sync()
    vol.synchronize()
    return vol


def mats():
    return [scenes.Material(albedo=(200, 150, 100)), scenes.Material(albedo=(255, 230, 180), emissive=True)]


def test_empty_map(gpu):
    nw, aw = empty_words((3, 2, 3))
    sc = Scene("empty", (3, 2, 3), nw, aw, mats(), cam_pos=(1.5, 1.0, 1.5), cam_orient=(20.0, 30.0, 0.0),
               width=40, height=24)
    o = oracle.load("f32")
    vol = load(sc)
    try:
        for f in range(2):
            nreq, nlit = run_frame_and_compare(vol, sc, o, f, 10.0 + f, sc.width, sc.height, mark_visible=1 << 20,
                                               full_draw=True)
            assert nreq == 0 and nlit == 0
        m, c, v = vol.export_state()
        assert not (m[:, 0] & 3).any()
    finally:
        vol.close()


def test_one_chunk_map(gpu):
    nw, aw = empty_words((1, 1, 1))
    rng = np.random.default_rng(3)
    for z in range(8):
        for x in range(8):
            h = int(rng.integers(1, 6))
            for y in range(h):
                solid(nw, aw, z, y, x, material=int(rng.integers(0, 2)) if y == h - 1 else 0,
                      normal=(rng.normal(), 1.0, rng.normal()))
    sc = Scene("one", (1, 1, 1), nw, aw, mats(), cam_pos=(-0.5, 1.2, -0.5), cam_orient=(30.0, 45.0, 0.0),
               width=32, height=32, params={"numDirectRays": 8})
    o = oracle.load("f32")
    vol = load(sc)
    try:
        lit = 0
        for f in range(3):
            _, n = run_frame_and_compare(vol, sc, o, f, 20.0 + f / 60.0, sc.width, sc.height, mark_visible=1 << 20,
                                         full_draw=True)
            lit += n
        assert lit > 3 * 64
    finally:
        vol.close()


def test_ragged_request_blocks(gpu):
    counts = (1, 31, 32, 33)
    nw, aw = empty_words((5, 1, 1))
    for ci, n in enumerate(counts):
        for k in range(n):              # one exposed layer: every voxel is stored
            solid(nw, aw, k // 8, 0, ci * 8 + k % 8)
    nw[:, :, 32:40] = pack_normal_word(np.uint32(0), np.array([0.0, 1.0, 0.0]))   # a solid chunk
    aw[:, :, 32:40] = (120 << 24) | (200 << 16) | (90 << 8)
    sc = Scene("ragged", (5, 1, 1), nw, aw, mats(), cam_pos=(2.5, 2.0, -1.5), cam_orient=(40.0, 0.0, 0.0),
               width=48, height=32, params={"numDirectRays": 4, "numSpecularRays": 15})
    o = oracle.load("f32")
    vol = load(sc)
    try:
        m, c, _ = vol.export_state()
        stored = [int(sum(bin(int(w)).count("1") for w in c[t, 7:23])) for t in range(5)]
        assert stored == [1, 31, 32, 33, 296]
        for f in range(3):
            nreq, nlit = run_frame_and_compare(vol, sc, o, f, 30.0 + f / 60.0, sc.width, sc.height,
                                               mark_visible=1 << 20)
            assert nreq == 1 + 1 + 1 + 2 + 10 and nlit == 1 + 31 + 32 + 33 + 296
    finally:
        vol.close()


@pytest.mark.parametrize("bad", ["samples", "bounces", "specular"])
def test_lighting_limits_refused(gpu, bad):
    sc = scenes.config_small(seed=41)
    vol = load(sc)
    try:
        vol.mark_visible(1 << 20)
        vol.sync()
        vol.synchronize()
        before = vol.export_state()
        errors = vol.lib.DNX_error_count()
        n = 1
        if bad == "samples":
            n = 5                                     # more than the 4 samples a pass holds
        elif bad == "bounces":
            vol.v.diffuseBounceLimit = 9              # more than 8
        else:
            vol.v.numSpecularRays = 0
        vol.update_lighting(n, 1000, 50.0)
        vol.synchronize()
        assert vol.lib.DNX_error_count() > errors
        for a, b in zip(before, vol.export_state()):
            assert np.array_equal(a, b)
    finally:
        vol.close()

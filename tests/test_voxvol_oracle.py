"""The oracle's .voxvol codec (voxel.c:300-654) pinned by the reference's own
saved file: decoding tests/golden/demo.voxvol (the reference's
assets/volumes/demo.voxvol, written by its DN_save_volume) and encoding the
result again reproduces the file byte for byte, which fixes the palette rule,
the palette order, the 255-voxel run limit and every field's width. Synthetic
chunks then cover what the demo file may not: palettes at the half-count
threshold, chunks of 0-3 voxels, full 512-voxel runs, unused slots."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import voxvol

DEMO = Path(__file__).resolve().parent / "golden" / "demo.voxvol"


@pytest.fixture(scope="module")
def demo():
    return DEMO.read_bytes()


def test_demo_round_trips_byte_for_byte(demo):
    vf = voxvol.decode(demo)
    assert voxvol.encode(vf) == demo


def test_demo_contents(demo):
    vf = voxvol.decode(demo)
    assert vf.map_size == (10, 3, 10)
    used = [c for c in vf.chunks if c.voxels is not None]
    assert len(vf.chunks) == 256 and len(used) == 240
    assert sum(c.num_voxels for c in used) == 28776
    mats = set()
    for c in used:
        solid = (c.voxels[:, 0] >> 24) != 255
        assert int(solid.sum()) == c.num_voxels
        assert not (c.voxels[solid, 1] & 0xFF).any()          # albedo's low byte is never stored
        mats |= set((c.voxels[solid, 0] >> 24).tolist())
    assert mats == {0, 1, 2, 3, 4}
    pos = [c.pos for c in used]
    assert len(set(pos)) == len(pos)                         # one slot per tile
    cam = vf.tail[:8]
    assert cam[6] == 90.0 and cam[7] == 0                     # fov, view mode


def solid_chunk(rng, n, distinct_normals, distinct_albedo, material=7):
    vox = np.zeros((512, 2), np.uint32)
    vox[:, 0] = 0xFFFFFFFF
    idx = np.sort(rng.choice(512, n, replace=False))
    nrm = rng.integers(0, 1 << 24, max(distinct_normals, 1)).astype(np.uint32)
    alb = (rng.integers(0, 1 << 24, max(distinct_albedo, 1)).astype(np.uint32) << 8)
    vox[idx, 0] = (np.uint32(material) << 24) | nrm[np.arange(n) % max(distinct_normals, 1)]
    vox[idx, 1] = alb[np.arange(n) % max(distinct_albedo, 1)]
    return voxvol.Chunk((1, 0, 2), vox, n)


@pytest.mark.parametrize("n,dn,da", [(0, 0, 0), (1, 1, 1), (2, 1, 1), (3, 1, 1), (4, 1, 2), (40, 19, 20),
                                     (40, 20, 19), (300, 149, 150), (512, 1, 1), (512, 255, 256)])
def test_chunk_round_trip_and_palette_rule(n, dn, da):
    rng = np.random.default_rng(n * 1000 + dn * 10 + da)
    ch = solid_chunk(rng, n, dn, da)
    rec = voxvol.encode_chunk(ch, (4, 4, 4))
    back, used = voxvol.decode_chunk(rec, (4, 4, 4))
    assert used == len(rec)
    np.testing.assert_array_equal(back.voxels, ch.voxels)
    assert back.num_voxels == n
    # a palette is written iff its distinct entries stay below half the voxel count
    nn = rec[12]
    distinct_n = len({int(w) & 0xFFFFFF for w in ch.voxels[:, 0] if (int(w) >> 24) != 255})
    assert nn == (distinct_n if distinct_n < n // 2 else 0)
    q = 13 + 3 * nn
    distinct_a = len({int(w) for w, m in zip(ch.voxels[:, 1], ch.voxels[:, 0]) if (int(m) >> 24) != 255})
    assert rec[q] == (distinct_a if distinct_a < n // 2 else 0)


def test_full_chunk_runs_are_split_at_255():
    ch = solid_chunk(np.random.default_rng(1), 512, 1, 1)
    rec = voxvol.encode_chunk(ch, (4, 4, 4))
    nn, = rec[12:13]
    q = 13 + 3 * nn
    na = rec[q]
    q += 1 + 3 * na
    runs = []
    while q < len(rec):
        runs.append(rec[q + 1])
        q += 2 + 2 * rec[q + 1]                              # palette indices: 1 byte each
    assert runs == [255, 255, 2]


def test_unused_slot_is_position_only():
    ch = voxvol.Chunk((-1, -1, -1))
    rec = voxvol.encode_chunk(ch, (4, 4, 4))
    assert len(rec) == 12
    back, used = voxvol.decode_chunk(rec, (4, 4, 4))
    assert back.voxels is None and used == 12

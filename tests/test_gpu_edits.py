"""Host voxel edits of the engine (DNX_set_voxels_dense, DN_shape_sphere) against
the oracle's dense edit model, bit for bit: every voxel word pair and every
tile's chunk-present flag after seeded random terrain and a run of seeded
sphere fills, carves and normal-flipping carves (integer centres, whose centre
voxel quantises a NaN normal, spheres over the map edge, radii 0.5 to 20).
The volume needs a HIP device, so this runs with the GPU tests."""
from __future__ import annotations

import numpy as np
import pytest

from doonengine_amd.volume import Volume
from oracle import load

pytestmark = pytest.mark.gpu

MAP = (6, 4, 5)


def spheres(rng, n):
    X, Y, Z = (8 * m for m in MAP)
    for k in range(n):
        integer = k % 5 == 0
        c = rng.integers(-6, [X + 6, Y + 6, Z + 6]).astype(float) if integer else \
            rng.uniform(-6, 6, 3) + rng.uniform(0, 1, 3) * (X, Y, Z)
        r = float(rng.uniform(0.5, 20.0))
        kind = rng.integers(0, 3)
        if kind == 0:
            yield 255, (0, 0, 0), tuple(c), r, False
        elif kind == 1:
            yield 255, tuple(int(v) for v in rng.integers(0, 256, 3)), tuple(c), r, True
        else:
            yield int(rng.integers(1, 200)), tuple(int(v) for v in rng.integers(0, 256, 3)), tuple(c), r, False


@pytest.mark.parametrize("seed", [0, 1])
def test_sphere_edits_match_oracle(gpu, seed):
    rng = np.random.default_rng(seed)
    o = load("f32")
    X, Y, Z = (8 * m for m in MAP)
    solid = rng.random((Z, Y, X)) < 0.3
    normal = np.where(solid, (rng.integers(1, 200, (Z, Y, X)).astype(np.uint32) << 24)
                      | rng.integers(0, 1 << 24, (Z, Y, X)).astype(np.uint32), np.uint32(0xFF000000))
    albedo = rng.integers(0, 1 << 32, (Z, Y, X), dtype=np.uint64).astype(np.uint32) & np.uint32(0xFFFFFF00)
    grid = o.edit_grid(MAP)
    grid.set_dense(normal, albedo)
    with Volume(MAP, min_chunks=16) as vol:
        vol.set_dense(normal, albedo)
        for k, (mat, alb, c, r, flip) in enumerate(spheres(rng, 60)):
            vol.shape_sphere(mat, alb, c, r, flip_normals=flip)
            grid.shape_sphere(mat, alb, c, r, flip_normals=flip)
            if k % 20 == 19:
                n, a = vol.get_dense((X, Y, Z))
                np.testing.assert_array_equal(n, grid.normal, err_msg=f"normal words after sphere {k}")
                np.testing.assert_array_equal(a, grid.albedo, err_msg=f"albedo words after sphere {k}")
        flags = np.array([[[vol.chunk_exists((x, y, z)) for x in range(MAP[0])] for y in range(MAP[1])]
                          for z in range(MAP[2])])
        np.testing.assert_array_equal(flags, grid.flag.astype(bool))
        assert 0 < flags.sum() < flags.size


def terrain(rng, upper_empty=True):
    X, Y, Z = (8 * m for m in MAP)
    solid = rng.random((Z, Y, X)) < 0.3
    if upper_empty:
        solid[:, Y // 2:, :] = False               # fills there add chunks during the batch
    normal = np.where(solid, (rng.integers(1, 200, (Z, Y, X)).astype(np.uint32) << 24)
                      | rng.integers(0, 1 << 24, (Z, Y, X)).astype(np.uint32), np.uint32(0xFF000000))
    albedo = rng.integers(0, 1 << 32, (Z, Y, X), dtype=np.uint64).astype(np.uint32) & np.uint32(0xFFFFFF00)
    return normal, albedo


def assert_same_host(a: Volume, b: Volume, what: str):
    ma, ca, na = a.host_state()
    mb, cb, nb = b.host_state()
    assert na == nb, f"nextChunk differs {what}"
    np.testing.assert_array_equal(ma["flag"], mb["flag"], err_msg=f"map flags {what}")
    live = ma["flag"] != 0
    np.testing.assert_array_equal(ma["chunkIndex"][live], mb["chunkIndex"][live], err_msg=f"chunk slots {what}")
    assert len(ca) == len(cb), f"chunk capacity differs {what}"
    for f in ("pos", "updated", "numVoxels", "numVoxelsGpu", "voxels"):
        np.testing.assert_array_equal(ca[f], cb[f], err_msg=f"chunk field {f} {what}")


@pytest.mark.parametrize("seed", [3, 4])
def test_batched_device_edits_match_sequential(gpu, seed):
    """DNX_shape_spheres (device, batched) against DN_shape_sphere one by one on
    an identical volume: the whole CPU chunk array (slots, pos, updated,
    counts, words), the map handles and nextChunk bit for bit after every batch
    (batches of 1 to 40 edits, chunk memory growing inside a batch, host-side
    DN_set_voxel / DN_remove_voxel edits between batches, which the device copy
    must pick up), then the uploaded device state after a sync."""
    rng = np.random.default_rng(seed)
    normal, albedo = terrain(rng)
    edits = list(spheres(rng, 90))
    with Volume(MAP, min_chunks=16) as seq, Volume(MAP, min_chunks=16) as bat:
        for v in (seq, bat):
            v.set_dense(normal, albedo)
        assert_same_host(seq, bat, "after terrain")
        k, sizes = 0, [1, 7, 40, 2, 25, 15]
        for bi, size in enumerate(sizes):
            chunk = edits[k:k + size]
            k += size
            for mat, alb, c, r, flip in chunk:
                seq.shape_sphere(mat, alb, c, r, flip_normals=flip)
            bat.shape_spheres(chunk)
            assert_same_host(seq, bat, f"after batch {bi}")
            # host-side edits between batches, on both volumes
            for j in range(6):
                mp = tuple(int(x) for x in rng.integers(0, MAP))
                cp = tuple(int(x) for x in rng.integers(0, 8, 3))
                if j % 3 == 2:
                    for v in (seq, bat):
                        v.remove_voxel(mp, cp) if v.chunk_exists(mp) else None
                else:
                    m = int(rng.integers(1, 200))
                    for v in (seq, bat):
                        v.set_voxel(mp, cp, m, (0.0, 1.0, 0.0), (9, 8, 7))
            assert_same_host(seq, bat, f"after host edits {bi}")
        assert k == len(edits)
        for v in (seq, bat):
            v.request_all()
            v.sync()
            v.synchronize()
        sa, sb = seq.export_state(), bat.export_state()
        for x, y, name in zip(sa, sb, ("map", "chunks", "voxels")):
            np.testing.assert_array_equal(x, y, err_msg=f"device {name} after sync")


def test_pick_ray_matches_oracle(gpu):
    """DN_step_map, the CPU pick ray (voxel.c:1195-1272), against the oracle's
    restatement on the same edited volume: hit flag, voxel position, face normal
    of the last step and the voxel read back, for random, axis-parallel, exact
    diagonal (tie) and outside-the-map rays with several step limits."""
    from oracle.pick import step_map
    rng = np.random.default_rng(9)
    normal, albedo = terrain(rng, upper_empty=False)
    o = load("f32")
    grid = o.edit_grid(MAP)
    grid.set_dense(normal, albedo)
    X, Y, Z = (8 * m for m in MAP)
    with Volume(MAP, min_chunks=16) as vol:
        vol.set_dense(normal, albedo)
        for mat, alb, c, r, flip in list(spheres(rng, 12)):
            vol.shape_sphere(mat, alb, c, r, flip_normals=flip)
            grid.shape_sphere(mat, alb, c, r, flip_normals=flip)
        rays = []
        for k in range(600):
            org = rng.uniform(-0.3, 1.3, 3) * np.array(MAP)
            if k % 4 == 0:
                d = np.zeros(3); d[rng.integers(0, 3)] = rng.choice([-1.0, 1.0])
            elif k % 4 == 1:
                d = rng.choice([-1.0, 1.0], 3)
                org = (np.floor(org * 8) + 0.5) / 8            # voxel centres: exact ties
            else:
                d = rng.normal(size=3)
            rays.append((d.astype(np.float32), org.astype(np.float32), int(rng.choice([1, 7, 40, 200]))))
        hits = 0
        for d, org, steps in rays:
            g = vol.pick(tuple(float(v) for v in d), tuple(float(v) for v in org), steps)
            hit, pos, n, idx = step_map(grid.normal, grid.flag, d, org, steps)
            assert g[0] == hit and g[2] == n, (d, org, steps)
            if hit:
                hits += 1
                assert g[1] == pos
                w, a = int(grid.normal[idx]), int(grid.albedo[idx])
                assert g[3].material == w >> 24
                assert (g[3].albedo.r, g[3].albedo.g, g[3].albedo.b) == ((a >> 24) & 255, (a >> 16) & 255, (a >> 8) & 255)
        assert hits > 100

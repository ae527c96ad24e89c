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

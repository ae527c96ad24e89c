"""The oracle's CPU pick ray (voxel.c:1195-1272) pinned by closed forms: an
axis-parallel ray stops at the first solid voxel on its line with the face
normal of the last step; a ray starting in a solid voxel hits it at once with
the untouched (-1000, -1000, -1000) normal; the strict-comparison tie rule steps
z, then y, then x along an exact diagonal (where the GPU's `<=` would step all
three at once); negative positions never index a chunk; max_steps bounds the
walk."""
from __future__ import annotations

import numpy as np

from oracle.pick import step_map

MAP = (2, 2, 2)


def grid():
    normal = np.full((16, 16, 16), 0xFF000000, np.uint32)
    flag = np.ones((2, 2, 2), np.uint8)
    return normal, flag


def test_axis_ray_stops_at_first_solid_voxel():
    normal, flag = grid()
    normal[5, 6, 11] = (3 << 24) | 0x7F7F7F          # (x=11, y=6, z=5)
    normal[5, 6, 13] = (4 << 24)
    hit, pos, n, idx = step_map(normal, flag, (1.0, 0.0, 0.0), ((2 + 0.5) / 8, (6 + 0.5) / 8, (5 + 0.5) / 8), 64)
    assert hit and pos == (11, 6, 5) and n == (-1, 0, 0) and idx == (5, 6, 11)
    hit, pos, n, _ = step_map(normal, flag, (-1.0, 0.0, 0.0), ((15 + 0.5) / 8, 6.5 / 8, 5.5 / 8), 64)
    assert hit and pos == (13, 6, 5) and n == (1, 0, 0)


def test_start_inside_solid_voxel():
    normal, flag = grid()
    normal[1, 2, 3] = 7 << 24
    hit, pos, n, _ = step_map(normal, flag, (0.3, 0.5, 0.8), (3.5 / 8, 2.5 / 8, 1.5 / 8), 10)
    assert hit and pos == (3, 2, 1) and n == (-1000, -1000, -1000)


def test_diagonal_ties_step_z_then_y_then_x():
    normal, flag = grid()
    seen = []
    for steps in range(1, 7):
        _, pos, n, _ = step_map(normal, flag, (1.0, 1.0, 1.0), (4.5 / 8, 4.5 / 8, 4.5 / 8), steps)
        seen.append(n)
    assert seen == [(0, 0, -1), (0, -1, 0), (-1, 0, 0)] * 2


def test_missing_chunk_and_negative_positions():
    normal, flag = grid()
    normal[2, 2, 2] = 5 << 24
    flag[0, 0, 0] = 0                                   # the voxel's chunk is absent: no hit there
    hit, _, _, _ = step_map(normal, flag, (1.0, 1e-4, 1e-4), (-3.0, 2.5 / 8, 2.5 / 8), 40)
    assert not hit
    flag[0, 0, 0] = 1
    hit, pos, n, _ = step_map(normal, flag, (1.0, 1e-4, 1e-4), (-3.0, 2.5 / 8, 2.5 / 8), 40)
    assert hit and pos == (2, 2, 2) and n == (-1, 0, 0)


def test_max_steps_bounds_the_walk():
    normal, flag = grid()
    normal[0, 0, 15] = 1 << 24
    hit, pos, _, _ = step_map(normal, flag, (1.0, 0.0, 0.0), (0.5 / 8, 0.5 / 8, 0.5 / 8), 14)
    assert not hit and pos == (14, 0, 0)
    hit, pos, _, _ = step_map(normal, flag, (1.0, 0.0, 0.0), (0.5 / 8, 0.5 / 8, 0.5 / 8), 16)
    assert hit and pos == (15, 0, 0)

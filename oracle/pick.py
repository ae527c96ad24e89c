"""The oracle's restatement of the reference's CPU pick ray, DN_step_map
(src/DoonEngine/voxel.c:1195-1272), over a dense voxel grid (test
infrastructure: only tests/ use it, as the checker).

A voxel-unit DDA with strict comparisons: when side distances tie, x loses to
y and z, and y loses to z, so z wins three-way ties (voxel.c:1241-1266) —
unlike the GPU DDA, whose `<=` mask steps every tied axis at once. The start
side distances are computed as in the C source, where the 0.5 literals are
double: sign * (pos - rayPos) in float, the rest in double, then rounded to
float (voxel.c:1213-1215). Negative positions also move the map position down
by one (voxel.c:1225-1230), so they never index a chunk.

Grid: `normal` (Z, Y, X) u32 normal words (material in the top byte, 255
empty) and `flag` (mz, my, mx) chunk-present flags, as oracle.EditGrid holds.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def _sign(v: float) -> int:
    return 1 if v > 0 else (-1 if v < 0 else 0)


def step_map(normal: np.ndarray, flag: np.ndarray, ray_dir, ray_pos, max_steps: int):
    """Returns (hit, hit_pos (x, y, z) voxel, hit_normal, voxel word index (z, y, x) or None)."""
    mz, my, mx = flag.shape
    hit_normal = (-1000, -1000, -1000)
    d = [f32(v) for v in ray_dir]
    p = [f32(v) * f32(8) for v in ray_pos]                      # rayPos * DN_CHUNK_SIZE
    inv = [f32(1) / v if v != 0 else (f32(np.inf) if math.copysign(1, v) > 0 else f32(-np.inf)) for v in d]
    sgn = [_sign(v) for v in d]
    pos = [int(math.floor(float(v))) for v in p]
    delta = [f32(abs(v)) for v in inv]
    side = []
    for a in range(3):
        first = f32(f32(sgn[a]) * f32(f32(pos[a]) - p[a]))      # float
        side.append(f32((float(first) + sgn[a] * 0.5 + 0.5) * float(delta[a])))
    for _ in range(max_steps):
        mp = [int(v / 8) if v >= 0 else -int(-v / 8) for v in pos]   # C division truncates
        cp = [v - 8 * m for v, m in zip(pos, mp)]                      # C remainder
        mp = [m - 1 if v < 0 else m for m, v in zip(mp, pos)]
        if 0 <= mp[0] < mx and 0 <= mp[1] < my and 0 <= mp[2] < mz and flag[mp[2], mp[1], mp[0]]:
            w = int(normal[pos[2], pos[1], pos[0]])
            if (w >> 24) != 255:
                return True, tuple(pos), hit_normal, (pos[2], pos[1], pos[0])
        if side[0] < side[1]:
            a = 0 if side[0] < side[2] else 2
        else:
            a = 1 if side[1] < side[2] else 2
        side[a] = f32(side[a] + delta[a])
        pos[a] += sgn[a]
        n = [0, 0, 0]
        n[a] = -sgn[a]
        hit_normal = tuple(n)
    return False, tuple(pos), hit_normal, None

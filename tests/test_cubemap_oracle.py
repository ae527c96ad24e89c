"""The oracle's sky cubemap sampler (texture(samplerCube, dir) with the
reference's GL_LINEAR / GL_CLAMP_TO_EDGE RGB8 cubemap, main.c:335-353)
pinned by closed forms: constant faces show which face a direction selects
(axes, off-axis directions, ties); faces whose texels encode their own column
and row show the (s, t) orientation of every face, stated as face axes
(normal N, s axis S, t axis T with S x T = -N, the cube map convention) and
independent of the oracle's per-face table; bilinear filtering reproduces a
linear ramp exactly between texel centres and clamps at the face edges."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import load

N = 16
# face normal, s axis, t axis for +X, -X, +Y, -Y, +Z, -Z
AXES = [((1, 0, 0), (0, 0, -1), (0, -1, 0)), ((-1, 0, 0), (0, 0, 1), (0, -1, 0)),
        ((0, 1, 0), (1, 0, 0), (0, 0, 1)), ((0, -1, 0), (1, 0, 0), (0, 0, -1)),
        ((0, 0, 1), (1, 0, 0), (0, -1, 0)), ((0, 0, -1), (-1, 0, 0), (0, -1, 0))]


@pytest.fixture(scope="module")
def o():
    return load("f32")


def scene_with(o, faces):
    s = o.scene((1, 1, 1), np.zeros((1, 3), np.uint32), np.zeros((1, 24), np.uint32), np.zeros((1, 4), np.uint32),
                np.zeros(256, dtype=[("p", "<u4", 8)]))
    o.set_cubemap(s, faces)
    return s


def test_face_axes_follow_the_cube_map_convention():
    for n, sx, tx in AXES:
        assert tuple(np.cross(sx, tx)) == tuple(-np.array(n))


def test_constant_faces_select_by_major_axis(o):
    faces = np.zeros((6, N, N, 3), np.uint8)
    for f in range(6):
        faces[f] = (10 * f + 5, 100 + f, 200 - f)
    s = scene_with(o, faces)
    for f, (n, sx, tx) in enumerate(AXES):
        for d in (np.array(n, float), np.array(n) + 0.3 * np.array(sx) - 0.6 * np.array(tx)):
            got = o.sky_color(s, d.astype(np.float32))
            # four equal texels under weights that sum to one: within an ulp of the texel
            np.testing.assert_allclose(got, faces[f, 0, 0].astype(np.float32) / np.float32(255), rtol=1.2e-7)
    # equal magnitudes: x before y before z
    assert abs(o.sky_color(s, np.array([1, 1, 0.5], np.float32))[0] * 255 - 5) < 1e-4
    assert abs(o.sky_color(s, np.array([-0.2, -1, 1], np.float32))[0] * 255 - 35) < 1e-4


def test_texel_coordinates_orient_every_face(o):
    faces = np.zeros((6, N, N, 3), np.uint8)
    j, i = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    faces[:, :, :, 0] = (i * 8)[None]
    faces[:, :, :, 1] = (j * 8)[None]
    s = scene_with(o, faces)
    rng = np.random.default_rng(2)
    for f, (n, sx, tx) in enumerate(AXES):
        for _ in range(20):
            e1, e2 = rng.uniform(-0.8, 0.8, 2)
            d = (np.array(n) + e1 * np.array(sx) + e2 * np.array(tx)).astype(np.float32)
            u = (e1 + 1) / 2 * N - 0.5
            v = (e2 + 1) / 2 * N - 0.5
            got = o.sky_color(s, d)
            np.testing.assert_allclose(got[0] * 255, 8 * np.clip(u, 0, N - 1), atol=2e-3)
            np.testing.assert_allclose(got[1] * 255, 8 * np.clip(v, 0, N - 1), atol=2e-3)


def test_bilinear_centres_midpoints_and_edges(o):
    faces = np.zeros((6, N, N, 3), np.uint8)
    vals = np.arange(N) * 16 + 7
    faces[4, :, :, 2] = vals[None, :]            # +Z face: blue ramps along s
    s = scene_with(o, faces)
    f32 = np.float32
    for k in (0, 5, N - 2):
        for frac, expect in ((0.5, vals[k]), (1.0, (vals[k] + vals[k + 1]) / 2)):
            sc = (k + frac) / N * 2 - 1                    # s = (k + frac) / N
            got = o.sky_color(s, np.array([sc, 0.1, 1.0], f32))
            assert abs(got[2] * 255 - expect) < 2e-3, (k, frac)
    # beyond the outermost texel centres the face edge texel holds
    assert abs(o.sky_color(s, np.array([0.999, 0, 1], f32))[2] * 255 - vals[-1]) < 1e-3
    assert abs(o.sky_color(s, np.array([-0.999, 0, 1], f32))[2] * 255 - vals[0]) < 1e-3

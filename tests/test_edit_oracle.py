"""The oracle's host-edit model (voxel.c:1126-1160 set/add/remove chunk,
voxelShapes.c:239-333 DN_shape_sphere) pinned by brute force: the voxels a
sphere writes are exactly the lattice points whose FP32 squared distance
(x*x + y*y) + z*z from the centre is below r*r, evaluated independently with
numpy float32; normals quantise fromCenter / max|fromCenter| (so one component
of every written normal sits at 0, 254 or 255); carving the same sphere restores the
cleared grid and removes every chunk; a flipped shell's normals point inwards;
solid spheres leave existing voxels alone; spheres past the map edge write only
in-bounds voxels."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import load

f32 = np.float32
MAP = (4, 3, 5)


@pytest.fixture(scope="module")
def o():
    return load("f32")


def lattice(map_size):
    mx, my, mz = map_size
    z, y, x = np.meshgrid(np.arange(8 * mz), np.arange(8 * my), np.arange(8 * mx), indexing="ij")
    return x, y, z


def inside(map_size, c, r, bound=None):
    x, y, z = lattice(map_size)
    fx = x.astype(f32) - f32(c[0])
    fy = y.astype(f32) - f32(c[1])
    fz = z.astype(f32) - f32(c[2])
    d2 = (fx * fx + fy * fy) + fz * fz
    lim = f32(r) * f32(r) if bound is None else bound
    return d2 < lim, fx, fy, fz


def quant(n):
    # quickmath's QM_MIN/QM_MAX (x < y ? x : y): a NaN component clamps to 1
    n = np.where(n < f32(1), n, f32(1))
    n = np.where(n > f32(-1), n, f32(-1))
    return ((np.trunc(n * f32(255)).astype(np.int64) + 255) // 2).astype(np.uint32)


def expected_fill(map_size, c, r, material, albedo):
    m, fx, fy, fz = inside(map_size, c, r)
    mx = np.maximum(np.maximum(np.abs(fx), np.abs(fy)), np.abs(fz))
    with np.errstate(divide="ignore", invalid="ignore"):   # the centre voxel: 0 * (1/0) = NaN
        s = f32(1) / mx
        word = (np.uint32(material) << 24) | (quant(fx * s) << 16) | (quant(fy * s) << 8) | quant(fz * s)
    normal = np.where(m, word, np.uint32(0xFFFFFFFF))
    alb = (albedo[0] << 24) | (albedo[1] << 16) | (albedo[2] << 8)
    return m, normal, np.where(m, np.uint32(alb), np.uint32(0))


def tiles_of(mask, map_size):
    mx, my, mz = map_size
    return mask.reshape(mz, 8, my, 8, mx, 8).any(axis=(1, 3, 5))


@pytest.mark.parametrize("c,r", [((13.3, 11.7, 19.2), 9.6), ((7.0, 8.0, 9.0), 5.0), ((16.5, 12.5, 20.5), 0.9)])
def test_fill_sphere_is_the_fp32_ball(o, c, r):
    g = o.edit_grid(MAP)
    g.shape_sphere(3, (200, 100, 50), c, r)
    m, normal, alb = expected_fill(MAP, c, r, 3, (200, 100, 50))
    assert m.sum() > 0
    np.testing.assert_array_equal(g.normal, normal)
    np.testing.assert_array_equal(g.albedo, alb)
    np.testing.assert_array_equal(g.flag.astype(bool), tiles_of(m, MAP))
    counts = m.reshape(MAP[2], 8, MAP[1], 8, MAP[0], 8).sum(axis=(1, 3, 5))
    np.testing.assert_array_equal(g.count, counts)
    w = g.normal[m]
    comps = np.stack([(w >> 16) & 255, (w >> 8) & 255, w & 255])
    # 1/max|fc| is rounded, so the dominant component lands at 254.99998 or 255
    assert np.all(np.any((comps == 0) | (comps >= 254), axis=0))


def test_carve_restores_cleared_grid(o):
    c, r = (13.3, 11.7, 19.2), 9.6
    g = o.edit_grid(MAP)
    g.shape_sphere(3, (200, 100, 50), c, r)
    g.shape_sphere(255, (0, 0, 0), c, r + 0.25)
    assert not g.flag.any() and not g.count.any()
    assert np.all(g.normal == 0xFFFFFFFF) and not g.albedo.any()


def test_partial_carve_writes_callers_empty_word_and_flips_shell(o):
    c = (15.2, 11.9, 19.6)
    g = o.edit_grid(MAP)
    g.shape_sphere(7, (10, 20, 30), c, 11.0)
    before = g.normal.copy()
    g.shape_sphere(255, (1, 2, 3), c, 6.0, flip_normals=True, normal=(0.5, -1.0, 0.0))
    hole, fx, fy, fz = inside(MAP, c, 6.0)
    ring = inside(MAP, c, 6.0, bound=f32(7.0) * f32(7.0))[0] & ~hole
    # every voxel inside the carve radius is the caller's empty voxel word
    empty = (np.uint32(255) << 24) | (quant(f32(0.5)) << 16) | (quant(f32(-1.0)) << 8) | quant(f32(0.0))
    assert np.all(g.normal[hole] == empty)
    assert np.all(g.albedo[hole] == (1 << 24) | (2 << 16) | (3 << 8))
    # the next shell keeps its material and turns its normals towards the centre
    solid_ring = ring & ((before >> 24) != 255)
    assert solid_ring.sum() > 100
    w = g.normal[solid_ring]
    assert np.all((w >> 24) == 7)
    n = np.stack([((w >> 16) & 255).astype(np.int64) * 2 - 255, ((w >> 8) & 255).astype(np.int64) * 2 - 255,
                  (w & 255).astype(np.int64) * 2 - 255], axis=1)
    fc = np.stack([fx[solid_ring], fy[solid_ring], fz[solid_ring]], axis=1)
    assert np.all(np.sum(n * fc, axis=1) < 0)
    # untouched beyond the shell
    far = ~inside(MAP, c, 6.0, bound=f32(7.0) * f32(7.0))[0]
    np.testing.assert_array_equal(g.normal[far], before[far])


def test_solid_sphere_keeps_existing_voxels(o):
    g = o.edit_grid(MAP)
    g.shape_sphere(4, (9, 9, 9), (10.0, 10.0, 10.0), 6.0)
    first = g.normal.copy()
    g.shape_sphere(5, (1, 1, 1), (14.0, 10.0, 10.0), 6.0)
    a = inside(MAP, (10.0, 10.0, 10.0), 6.0)[0]
    b = inside(MAP, (14.0, 10.0, 10.0), 6.0)[0]
    np.testing.assert_array_equal(g.normal[a], first[a])
    assert np.all((g.normal[b & ~a] >> 24) == 5)


def test_sphere_past_map_edge(o):
    g = o.edit_grid(MAP)
    c, r = (-3.0, 2.5, 38.0), 7.5
    g.shape_sphere(2, (5, 6, 7), c, r)
    m, normal, _ = expected_fill(MAP, c, r, 2, (5, 6, 7))
    assert m.sum() > 0
    np.testing.assert_array_equal(g.normal, normal)


def test_set_word_counts_and_removal(o):
    g = o.edit_grid(MAP)
    g.set_word(1, 2, 3, (9 << 24) | 0x7F7F7F, 0x11223300)
    g.set_word(2, 2, 3, (9 << 24) | 0x7F7F7F, 0x11223300)
    assert g.flag[0, 0, 0] == 1 and g.count[0, 0, 0] == 2
    g.set_word(9, 9, 9, 0xFF7F7F7F, 0)          # empty into a missing chunk: no-op
    assert g.flag[1, 1, 1] == 0 and g.normal[9, 9, 9] == 0xFFFFFFFF
    g.set_word(1, 2, 3, 0xFF000000, 0)
    assert g.count[0, 0, 0] == 1 and g.normal[3, 2, 1] == 0xFF000000
    g.set_word(2, 2, 3, 0xFF000000, 0)           # last voxel: chunk removed and cleared
    assert g.flag[0, 0, 0] == 0 and g.count[0, 0, 0] == 0
    assert np.all(g.normal[:8, :8, :8] == 0xFFFFFFFF)

"""DN_shape_sphere with the reference's full signature (voxelShapes.h:40):
orientation, transform callback and user data. A C callback (compiled from
tests/transform_callbacks.c) is given, as the same function pointer, to the
library and to the oracle's sphere edit; every voxel word and chunk-present
flag must agree bit for bit, the callback must see the sphere's bounding box,
the user pointer and the inverse of translate(c) * rotation(orient), and the
identity callback must reproduce the edit without one. The volume needs a HIP
device, so this runs with the GPU tests."""
from __future__ import annotations

import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from doonengine_amd import _lib
from doonengine_amd.volume import Volume
from oracle import load

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
MAP = (5, 4, 5)


class CallbackLog(ctypes.Structure):
    _fields_ = [("calls", ctypes.c_uint32), ("shapeMin", ctypes.c_float * 3), ("shapeMax", ctypes.c_float * 3),
                ("inv", ctypes.c_float * 16)]


@pytest.fixture(scope="module")
def callbacks(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("needs gcc to build the test callbacks")
    out = tmp_path_factory.mktemp("cb") / "libtransform_callbacks.so"
    subprocess.run(["gcc", "-O1", "-shared", "-fPIC", "-std=c11", "-ffp-contract=off", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "transform_callbacks.c"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    return {name: ctypes.cast(getattr(lib, name), ctypes.c_void_p).value for name in ("pattern", "identity")}


def quat(axis, angle):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    s = np.sin(angle / 2)
    return tuple(float(v) for v in (*(axis * s), np.cos(angle / 2)))


def terrain(rng):
    X, Y, Z = (8 * m for m in MAP)
    solid = rng.random((Z, Y, X)) < 0.35
    solid[:, Y // 2:, :] = False
    normal = np.where(solid, (rng.integers(1, 200, (Z, Y, X)).astype(np.uint32) << 24)
                      | rng.integers(0, 1 << 24, (Z, Y, X)).astype(np.uint32), np.uint32(0xFF000000))
    albedo = rng.integers(0, 1 << 32, (Z, Y, X), dtype=np.uint64).astype(np.uint32) & np.uint32(0xFFFFFF00)
    return normal, albedo


def edits(rng, n):
    X, Y, Z = (8 * m for m in MAP)
    for k in range(n):
        c = tuple(float(v) for v in rng.uniform(-4, 4, 3) + rng.uniform(0, 1, 3) * (X, Y, Z))
        r = float(rng.uniform(1.5, 12.0))
        q = quat(rng.normal(size=3), float(rng.uniform(0, 2 * np.pi)))
        kind = k % 3
        if kind == 0:
            yield int(rng.integers(1, 200)), tuple(int(v) for v in rng.integers(0, 256, 3)), c, r, False, q
        elif kind == 1:
            yield 255, (0, 0, 0), c, r, True, q
        else:
            yield 255, (0, 0, 0), c, r, False, q


@pytest.mark.parametrize("seed", [0, 1])
def test_transform_callback_matches_oracle(gpu, callbacks, seed):
    rng = np.random.default_rng(100 + seed)
    o = load("f32")
    X, Y, Z = (8 * m for m in MAP)
    normal, albedo = terrain(rng)
    grid = o.edit_grid(MAP)
    grid.set_dense(normal, albedo)
    log_lib, log_or = CallbackLog(), CallbackLog()
    fn = callbacks["pattern"]
    with Volume(MAP, min_chunks=16) as vol:
        vol.set_dense(normal, albedo)
        for k, (mat, alb, c, r, flip, q) in enumerate(edits(rng, 40)):
            n0 = log_lib.calls
            vol.shape_sphere(mat, alb, c, r, flip_normals=flip, orient=q, func=fn,
                             user_data=ctypes.addressof(log_lib))
            grid.shape_sphere(mat, alb, c, r, flip_normals=flip, orient=q, fn=fn, user_data=ctypes.addressof(log_or))
            assert log_lib.calls == log_or.calls
            if log_lib.calls > n0:
                # bounding box and inverse transform seen by the callback
                assert tuple(log_lib.shapeMin) == tuple(np.float32(c) - np.float32(r))
                assert tuple(log_lib.shapeMax) == tuple(np.float32(c) + np.float32(r))
                assert np.array_equal(np.array(log_lib.inv, np.float32), np.array(log_or.inv, np.float32))
                x, y, z, w = q
                rot = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                                [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                                [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
                m = np.eye(4)
                m[:3, :3] = rot.T            # QuickMath's column-major m[col][row] as a row-major matrix
                m[:3, 3] = c
                inv = np.linalg.inv(m)
                got = np.array(log_lib.inv, np.float64).reshape(4, 4).T
                np.testing.assert_allclose(got, inv, atol=2e-5 * max(1.0, float(np.abs(c).max())))
            if k % 10 == 9:
                nn, aa = vol.get_dense((X, Y, Z))
                np.testing.assert_array_equal(nn, grid.normal, err_msg=f"normal words after edit {k}")
                np.testing.assert_array_equal(aa, grid.albedo, err_msg=f"albedo words after edit {k}")
        assert log_lib.calls > 1000
        flags = np.array([[[vol.chunk_exists((x, y, z)) for x in range(MAP[0])] for y in range(MAP[1])]
                          for z in range(MAP[2])])
        np.testing.assert_array_equal(flags, grid.flag.astype(bool))


def test_identity_callback_equals_no_callback(gpu, callbacks):
    rng = np.random.default_rng(7)
    normal, albedo = terrain(rng)
    X, Y, Z = (8 * m for m in MAP)
    with Volume(MAP, min_chunks=16) as a, Volume(MAP, min_chunks=16) as b:
        a.set_dense(normal, albedo)
        b.set_dense(normal, albedo)
        for mat, alb, c, r, flip, q in edits(rng, 30):
            a.shape_sphere(mat, alb, c, r, flip_normals=flip, orient=q, func=callbacks["identity"])
            b.shape_sphere(mat, alb, c, r, flip_normals=flip)
        for x, y in zip(a.get_dense((X, Y, Z)), b.get_dense((X, Y, Z))):
            np.testing.assert_array_equal(x, y)

"""Host-side checks that need no GPU: the C-ABI library loads and exports every
declared symbol, the bindings' struct layouts match, and the pure-CPU parts of
the API behave as the reference's C code (src/DoonEngine/voxel.c:1277-1316,
788-853) and QuickMath define them."""
from __future__ import annotations

import ctypes
import json
import re
import subprocess
import sys
from pathlib import Path

import numpy as np

import oracle
from doonengine_amd import _lib, scenes

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "doon" / "doon.h"


def lib():
    return _lib.load()


def test_library_exports_every_declared_symbol():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(DNX?_[A-Za-z0-9_]+)\s*\(", text))
    names -= {"DN_FLATTEN_INDEX"}
    names -= set(re.findall(r"static inline \w+\s+(DNX?_[A-Za-z0-9_]+)\s*\(", text))   # header-only helpers
    assert len(names) > 40
    L = lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the message callback global of globals.h:63
    ctypes.c_void_p.in_dll(L, "g_DN_message_callback")


def test_struct_layouts_match():
    _lib.check_layout(lib())
    assert ctypes.sizeof(_lib.DNmaterial) == 32


def test_compress_decompress_semantics():
    L = lib()
    rng = np.random.default_rng(1)
    for _ in range(300):
        n = rng.uniform(-1.3, 1.3, 3).astype(np.float32)
        mat = int(rng.integers(0, 256))
        alb = [int(a) for a in rng.integers(0, 256, 3)]
        vox = _lib.DNvoxel(mat, _lib.DNvec3(*[float(x) for x in n]), _lib.DNcolor(*alb))
        cv = L.DN_compress_voxel(vox)
        nc = np.clip(n, -1, 1)
        q = [(int(np.float32(x) * np.float32(255.0)) + 255) // 2 for x in nc]  # (int) truncates toward zero
        assert cv.normal == (mat << 24) | (q[0] << 16) | (q[1] << 8) | q[2]
        assert cv.albedo == (alb[0] << 24) | (alb[1] << 16) | (alb[2] << 8)
        dv = L.DN_decompress_voxel(cv)
        assert dv.material == mat and (dv.albedo.r, dv.albedo.g, dv.albedo.b) == tuple(alb)
        # the CPU decoder: (2v - 255) in integers, times 1.0f/255.0f (voxel.c:1307-1311)
        for k, b in enumerate(q):
            exp = np.float32(2 * b - 255) * np.float32(1.0 / 255.0)
            assert np.float32(getattr(dv.normal, "xyz"[k])) == exp


def test_separate_position_and_bounds():
    L = lib()
    for p in [(0, 0, 0), (7, 8, 9), (63, 17, 100), (-1, -8, -9)]:
        mp, cp = _lib.DNivec3(), _lib.DNivec3()
        L.DN_separate_position(_lib.DNivec3(*p), ctypes.byref(mp), ctypes.byref(cp))
        for k, a in enumerate(p):
            q = int(a / 8)                         # C division truncates toward zero
            assert getattr(mp, "xyz"[k]) == q and getattr(cp, "xyz"[k]) == a - 8 * q
    assert L.DN_in_chunk_bounds(_lib.DNivec3(7, 0, 3))
    assert not L.DN_in_chunk_bounds(_lib.DNivec3(8, 0, 3))
    assert not L.DN_in_chunk_bounds(_lib.DNivec3(0, -1, 3))


def _rot(angles):
    x, y, z = np.radians(np.asarray(angles, np.float64))
    sx, cx, sy, cy, sz, cz = np.sin(x), np.cos(x), np.sin(y), np.cos(y), np.sin(z), np.cos(z)
    m = np.zeros((3, 3))                       # m[col][row], quickmath.h:1214-1241
    m[0] = [cy * cz, cy * sz, -sy]
    m[1] = [sx * sy * cz - cx * sz, sx * sy * sz + cx * cz, sx * cy]
    m[2] = [cx * sy * cz + sx * sz, cx * sy * sz - sx * cz, cx * cy]
    return m


def test_cam_dir_closed_form():
    L = lib()
    for ang in [(0, 0, 0), (30, 45, 0), (-20, 170, 10), (89, -45, 5)]:
        d = L.DN_cam_dir(_lib.DNvec3(*ang))
        exp = _rot(ang)[2]                      # rotate_euler * (0, 0, 1)
        np.testing.assert_allclose([d.x, d.y, d.z], exp, atol=2e-6)


def test_view_projection_and_draw_matrices():
    L = lib()
    v = _lib.DNvolume()
    v.camPos = _lib.DNvec3(1.5, 2.0, -3.0)
    v.camOrient = _lib.DNvec3(25.0, 60.0, 0.0)
    v.camFOV = 90.0
    for aspect in (0.5625, 1.0, 1.5):
        view, proj = _lib.DNmat4(), _lib.DNmat4()
        L.DN_set_view_projection_matrices(ctypes.byref(v), aspect, 0.1, 100.0, ctypes.byref(view), ctypes.byref(proj))
        V = np.array(view.m, np.float64)        # [col][row]
        P = np.array(proj.m, np.float64)
        # the camera looks along DN_cam_dir: the view matrix maps cam + front to -z
        front = _rot((25.0, 60.0, 0.0))[2]
        pt = np.array([1.5, 2.0, -3.0]) + 2 * front
        vp = V.T @ np.append(pt, 1.0)
        np.testing.assert_allclose(vp[:3] / np.linalg.norm(vp[:3]), [0, 0, -1], atol=1e-5)
        iv, ivp = _lib.DNmat4(), _lib.DNmat4()
        L.DNX_draw_matrices(view, proj, ctypes.byref(iv), ctypes.byref(ivp))
        IV = np.array(iv.m, np.float64)
        np.testing.assert_allclose(IV.T @ V.T, np.eye(4), atol=2e-5)
        C = V.copy()
        C[3, :3] = 0
        exp = np.linalg.inv(C.T) @ np.linalg.inv(P.T)       # row-major product
        np.testing.assert_allclose(np.array(ivp.m).T, exp, rtol=2e-5, atol=2e-6)
        # perspective: quickmath.h:1264-1282 with aspect = 1/aspectRatio
        t = np.tan(np.radians(45.0))
        np.testing.assert_allclose([P[0, 0], P[1, 1], P[2, 3]], [1 / ((1 / aspect) * t), 1 / t, -1.0], rtol=1e-6)


def test_fixed_ray_tables():
    L = lib()
    v = _lib.DNvolume()
    ref = np.array(json.loads((ROOT / "tests" / "golden" / "sphere_points_15.json").read_text())["points"], np.float32)
    out = np.zeros((64, 3), np.float32)
    v.numSpecularRays = 15
    n = L.DNX_get_ray_points(ctypes.byref(v), 0, out.ctypes.data, 64)
    assert n == 15 and np.array_equal(out[:15], ref)           # the reference's table, value for value
    for k in (1, 8, 16, 32, 64):
        v.numDirectRays = k
        n = L.DNX_get_ray_points(ctypes.byref(v), 1, out.ctypes.data, 64)
        assert n == k
        np.testing.assert_allclose(out[:k], oracle.golden_points(k), atol=1e-7)


def test_scene_generators_deterministic():
    a = scenes.config_small()
    b = scenes.config_small()
    assert np.array_equal(a.normal_words, b.normal_words) and np.array_equal(a.albedo_words, b.albedo_words)
    c1 = scenes.config1()
    assert abs(c1.occupancy - 0.35) < 0.01
    assert c1.normal_words.shape == (128, 128, 128)
    em = ((c1.normal_words >> 24) == 3).sum() / max(((c1.normal_words >> 24) != 255).sum(), 1)
    assert 0.002 < em < 0.01                                  # ~0.5% emissive


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr

"""The oracle: CPU restatement of the reference's voxel pipeline (test infrastructure).

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this
package, and only as the checker. ``f32`` is the strict-IEEE FP32 build (the
bit-exact pin of the HIP kernels), ``f64`` the same algorithm in FP64.

Every function takes the device state in the reference's buffer layouts
(voxelShared.comp:19-95): map (T, 3) u32 [flags, lastUsed, voxelIndex],
chunks (T, 24) u32 [pos(3), samples, partialCounts(3), bitMask(16), pad],
voxels (N, 4) u32, materials (256,) structured.
"""
from __future__ import annotations

import ctypes
import time as _time
from ctypes import POINTER, Structure, c_float, c_int, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

class ORcounters(Structure):
    _fields_ = [("calls", c_uint64), ("sMap", c_uint64), ("sVox", c_uint64), ("hits", c_uint64)]


class ORscene(Structure):
    _fields_ = [("map", c_void_p), ("chunks", c_void_p), ("voxels", c_void_p), ("mats", c_void_p),
                ("mapSize", c_uint32 * 3), ("sunStrength", c_float * 3), ("skyBot", c_float * 3),
                ("skyTop", c_float * 3), ("flagsOut", c_void_p), ("lastUsedOut", c_void_p), ("cube", c_void_p),
                ("cubeSize", c_uint32)]


class ORdrawParams(Structure):
    _fields_ = [("invView", c_float * 16), ("invViewProj", c_float * 16), ("width", c_uint32), ("height", c_uint32),
                ("viewMode", c_uint32)]


class ORlightParams(Structure):
    _fields_ = [("time", c_float), ("numDiffuseSamples", c_uint32), ("maxDiffuseSamples", c_uint32),
                ("diffuseBounceLimit", c_uint32), ("specBounceLimit", c_uint32), ("sunDir", c_float * 3),
                ("ambient", c_float * 3), ("shadowSoftness", c_float), ("camPos", c_float * 3),
                ("numSpecRays", c_uint32), ("numDirectRays", c_uint32), ("specPoints", c_void_p),
                ("directPoints", c_void_p)]


class ORgrid(Structure):
    _fields_ = [("mx", ctypes.c_int32), ("my", ctypes.c_int32), ("mz", ctypes.c_int32), ("normal", c_void_p),
                ("albedo", c_void_p), ("flag", c_void_p), ("count", c_void_p)]


class EditGrid:
    """Dense host-edit model (voxel.c:1126-1160, voxelShapes.c:239-333): words
    (Z, Y, X) per voxel, chunk-present flag and stored-voxel count per tile."""

    def __init__(self, oracle: "Oracle", map_size):
        mx, my, mz = (int(v) for v in map_size)
        self.o = oracle
        self.normal = np.full((8 * mz, 8 * my, 8 * mx), 0xFFFFFFFF, np.uint32)
        self.albedo = np.zeros((8 * mz, 8 * my, 8 * mx), np.uint32)
        self.flag = np.zeros((mz, my, mx), np.uint8)
        self.count = np.zeros((mz, my, mx), np.int32)
        self.g = ORgrid(mx, my, mz, self.normal.ctypes.data, self.albedo.ctypes.data, self.flag.ctypes.data,
                        self.count.ctypes.data)

    def set_word(self, x: int, y: int, z: int, normal: int, albedo: int) -> None:
        self.o.lib.or_set_voxel_word(ctypes.byref(self.g), int(x), int(y), int(z), c_uint32(normal), c_uint32(albedo))

    def set_dense(self, normal_words: np.ndarray, albedo_words: np.ndarray) -> None:
        """DNX_set_voxels_dense from the origin: voxel by voxel, x fastest."""
        zs, ys, xs = normal_words.shape
        n = np.ascontiguousarray(normal_words, np.uint32)
        a = np.ascontiguousarray(albedo_words, np.uint32)
        self.o.lib.or_set_dense(ctypes.byref(self.g), 0, 0, 0, xs, ys, zs, n.ctypes.data_as(c_void_p),
                                a.ctypes.data_as(c_void_p))

    def shape_sphere(self, material: int, albedo, center, radius: float, flip_normals=False, normal=(0, 0, 0),
                     orient=None, fn=None, user_data=None):
        """DN_shape_sphere; `fn` is the address of a C VoxelTransformFunc (or None),
        called with `user_data` (an address) and the inverse of
        translate(center) * rotation(orient), orient an (x, y, z, w) quaternion."""
        n = np.asarray(normal, np.float32)
        word = (int(albedo[0]) << 24) | (int(albedo[1]) << 16) | (int(albedo[2]) << 8)
        q = None if orient is None else np.ascontiguousarray(orient, np.float32)
        self.o.lib.or_shape_sphere_fn(ctypes.byref(self.g), c_uint32(material), n.ctypes.data_as(c_void_p),
                                      c_uint32(word), c_int(bool(flip_normals)), c_float(center[0]),
                                      c_float(center[1]), c_float(center[2]), c_float(radius),
                                      None if q is None else q.ctypes.data_as(c_void_p),
                                      c_void_p(fn) if fn else None, c_void_p(user_data) if user_data else None)


RAYHIT_DTYPE = np.dtype([("hit", "<u4"), ("voxelIndex", "<u4"), ("mapSteps", "<u4"), ("voxelSteps", "<u4"),
                         ("pos", "<f4", 3), ("normal", "<f4", 3)])
PIXEL_DTYPE = np.dtype([("voxelIndex", "<u4"), ("mapSteps", "<u4"), ("voxelSteps", "<u4"), ("pad", "<u4"),
                        ("face", "<f4", 3)])
MATERIAL_DTYPE = np.dtype([("padding", "<f4", 2), ("emissive", "<u4"), ("opacity", "<f4"), ("refractIndex", "<f4"),
                           ("specular", "<f4"), ("reflectType", "<u4"), ("shininess", "<u4")])


class Oracle:
    def __init__(self, path: Path):
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `python oracle/build.py`")
        self.lib = ctypes.CDLL(str(path))
        self.lib.or_sizeof.restype = c_uint32
        self.real = np.float64 if self.lib.or_sizeof(7) == 8 else np.float32
        for name, res in (("or_rank", c_uint32), ("or_select", c_int), ("or_chunk_to_gpu", c_uint32),
                          ("or_visible_requests", c_uint32), ("or_rand", c_float)):
            getattr(self.lib, name).restype = res
        self.lib.or_rand.argtypes = [c_float]
        self.lib.or_rand_unit_sphere.argtypes = [c_float, c_void_p]
        assert self.lib.or_sizeof(0) == 32
        assert self.lib.or_sizeof(1) == ctypes.sizeof(ORscene)
        assert self.lib.or_sizeof(4) == ctypes.sizeof(ORdrawParams)
        assert self.lib.or_sizeof(6) == ctypes.sizeof(ORlightParams)

    # -- scene ------------------------------------------------------------
    def scene(self, map_size, map_, chunks, voxels, mats, sun=(0.6, 0.6, 0.6), sky_bot=(0.71, 0.85, 0.90),
              sky_top=(0.00, 0.45, 0.74)):
        keep = [np.ascontiguousarray(map_, np.uint32), np.ascontiguousarray(chunks, np.uint32),
                np.ascontiguousarray(voxels, np.uint32), np.ascontiguousarray(mats)]
        s = ORscene()
        s.map, s.chunks, s.voxels, s.mats = (k.ctypes.data for k in keep)
        s.mapSize[:] = [int(v) for v in map_size]
        s.sunStrength[:] = sun
        s.skyBot[:] = sky_bot
        s.skyTop[:] = sky_top
        s.flagsOut = None
        s.lastUsedOut = None
        s.cube = None
        s.cubeSize = 0
        s._keep = keep
        return s

    def set_cubemap(self, scene: ORscene, faces) -> None:
        """Sky cubemap of a scene: faces (6, n, n, 3) uint8 in GL order, None for the gradient."""
        if faces is None:
            scene.cube, scene.cubeSize = None, 0
            return
        f = np.ascontiguousarray(faces, np.uint8)
        scene._cube = f
        scene.cube, scene.cubeSize = f.ctypes.data, f.shape[1]

    def sky_color(self, scene: ORscene, direction) -> np.ndarray:
        d = np.ascontiguousarray(direction, np.float32)
        out = np.zeros(3, self.real)
        self.lib.or_sky_color(ctypes.byref(scene), d.ctypes.data_as(c_void_p), out.ctypes.data_as(c_void_p))
        return out

    # -- primitives -------------------------------------------------------
    def rank(self, chunk: np.ndarray, x: int, y: int, z: int) -> int:
        c = np.ascontiguousarray(chunk, np.uint32)
        return self.lib.or_rank(c.ctypes.data_as(c_void_p), c_uint32(x), c_uint32(y), c_uint32(z))

    def select(self, chunk: np.ndarray, n: int) -> int:
        c = np.ascontiguousarray(chunk, np.uint32)
        return self.lib.or_select(c.ctypes.data_as(c_void_p), c_uint32(n))

    def decompress(self, words) -> np.ndarray:
        w = np.ascontiguousarray(words, np.uint32)
        out = np.zeros(13, self.real)
        self.lib.or_decompress(w.ctypes.data_as(c_void_p), out.ctypes.data_as(c_void_p))
        return out

    def set_fresh_output(self, out: np.ndarray | None) -> None:
        """With `out` ((>= numRequests*32, 3) array of this oracle's real type,
        kept alive by the caller), lighting() also writes each voxel's diffuse
        samples of the pass, averaged, before the running mean. None disables."""
        self._fresh = out
        self.lib.or_set_fresh_output(None if out is None else out.ctypes.data_as(c_void_p))

    def set_rand_override(self, seeds=None, values=None) -> None:
        """Use `values` as the hash of exactly these `seeds` (tests only); None clears."""
        if seeds is None:
            self._ov = None
            self.lib.or_set_rand_override(None, None, c_uint32(0))
            return
        order = np.argsort(np.asarray(seeds, np.float32), kind="stable")
        s = np.ascontiguousarray(np.asarray(seeds, np.float32)[order])
        v = np.ascontiguousarray(np.asarray(values, np.float32)[order])
        self._ov = (s, v)
        self.lib.or_set_rand_override(s.ctypes.data_as(c_void_p), v.ctypes.data_as(c_void_p), c_uint32(len(s)))

    def rand(self, seed: float) -> float:
        return self.lib.or_rand(float(seed))

    def rand_unit_sphere(self, seed: float) -> np.ndarray:
        out = np.zeros(3, np.float32)
        self.lib.or_rand_unit_sphere(c_float(seed), out.ctypes.data_as(c_void_p))
        return out

    def edit_grid(self, map_size) -> EditGrid:
        return EditGrid(self, map_size)

    def chunk_to_gpu(self, mats: np.ndarray, cpu_voxels: np.ndarray):
        """cpu_voxels: (8, 8, 8, 2) u32 indexed [x][y][z] -> (chunk (24,), voxels (n, 4))."""
        cv = np.ascontiguousarray(cpu_voxels, np.uint32)
        m = np.ascontiguousarray(mats)
        ch = np.zeros(24, np.uint32)
        vx = np.zeros((512, 4), np.uint32)
        n = self.lib.or_chunk_to_gpu(m.ctypes.data_as(c_void_p), cv.ctypes.data_as(c_void_p),
                                     ch.ctypes.data_as(c_void_p), vx.ctypes.data_as(c_void_p))
        return ch, vx[:n].copy()

    def visible_requests(self, map_, chunks, fresh=None, split=1, frame=0) -> np.ndarray:
        m = np.ascontiguousarray(map_, np.uint32)
        c = np.ascontiguousarray(chunks, np.uint32)
        T = m.shape[0]
        out = np.zeros(T * 16 + 1, np.uint32)
        fr = None if fresh is None else np.ascontiguousarray(fresh, np.uint8)
        n = self.lib.or_visible_requests(m.ctypes.data_as(c_void_p), c.ctypes.data_as(c_void_p), c_uint32(T),
                                         None if fr is None else fr.ctypes.data_as(c_void_p), c_uint32(split),
                                         c_uint32(frame), out.ctypes.data_as(c_void_p))
        return out[:n].copy()

    def cast_rays(self, scene: ORscene, origins, dirs, ignore_first=False):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros(o.shape[0], RAYHIT_DTYPE)
        cnt = ORcounters()
        self.lib.or_cast_rays(ctypes.byref(scene), c_uint32(o.shape[0]), o.ctypes.data_as(c_void_p),
                              d.ctypes.data_as(c_void_p), c_uint32(int(ignore_first)), out.ctypes.data_as(c_void_p),
                              ctypes.byref(cnt))
        return out, {f: getattr(cnt, f) for f, _ in cnt._fields_}

    def draw(self, scene: ORscene, inv_view: np.ndarray, inv_view_proj: np.ndarray, width: int, height: int,
             view_mode: int = 0, flags_out: np.ndarray | None = None, threads: int = 1):
        """inv_view / inv_view_proj: (4,4) float32 indexed [col, row]. threads > 1
        splits the rows over host threads (same result)."""
        dp = ORdrawParams()
        dp.invView[:] = [float(v) for v in np.asarray(inv_view, np.float32).reshape(-1)]
        dp.invViewProj[:] = [float(v) for v in np.asarray(inv_view_proj, np.float32).reshape(-1)]
        dp.width, dp.height, dp.viewMode = width, height, view_mode
        rgba = np.zeros((height, width, 4), np.float32)
        px = np.zeros((height, width), PIXEL_DTYPE)
        cnt = ORcounters()
        fo = None if flags_out is None else flags_out.ctypes.data_as(c_void_p)
        t0 = _time.perf_counter()
        if threads > 1:
            self.lib.or_draw_mt(ctypes.byref(scene), ctypes.byref(dp), rgba.ctypes.data_as(c_void_p),
                                px.ctypes.data_as(c_void_p), fo, ctypes.byref(cnt), c_uint32(threads))
        else:
            self.lib.or_draw(ctypes.byref(scene), ctypes.byref(dp), rgba.ctypes.data_as(c_void_p),
                             px.ctypes.data_as(c_void_p), fo, ctypes.byref(cnt))
        dt = _time.perf_counter() - t0
        return rgba, px, dict({f: getattr(cnt, f) for f, _ in cnt._fields_}, seconds=dt)

    def lighting(self, scene: ORscene, requests: np.ndarray, *, time: float, num_diffuse=1, max_diffuse=1000,
                 diffuse_bounces=5, spec_bounces=2, sun_dir=(1, 1, 1), ambient=(0.01, 0.01, 0.01), softness=10.0,
                 cam_pos=(0, 0, 0), spec_points: np.ndarray, direct_points: np.ndarray | None = None, threads: int = 1):
        """One lighting dispatch. Returns (voxels_out, map_out, chunks_out, prequant, vox_of, counters).
        threads > 1 splits the requests over host threads (same result)."""
        s = scene
        m_in, c_in, v_in = s._keep[0], s._keep[1], s._keep[2]
        lp = ORlightParams()
        lp.time = time
        lp.numDiffuseSamples, lp.maxDiffuseSamples = num_diffuse, max_diffuse
        lp.diffuseBounceLimit, lp.specBounceLimit = diffuse_bounces, spec_bounces
        sd = np.asarray(sun_dir, np.float32)
        lp.sunDir[:] = [float(v) for v in sd]
        lp.ambient[:] = [float(v) for v in ambient]
        lp.shadowSoftness = softness
        lp.camPos[:] = [float(v) for v in cam_pos]
        sp = np.ascontiguousarray(spec_points, np.float32).reshape(-1, 3)
        dp = np.zeros((1, 3), np.float32) if direct_points is None else np.ascontiguousarray(direct_points,
                                                                                             np.float32).reshape(-1, 3)
        lp.numSpecRays = sp.shape[0]
        lp.numDirectRays = 0 if direct_points is None else dp.shape[0]
        lp.specPoints = sp.ctypes.data
        lp.directPoints = dp.ctypes.data
        req = np.ascontiguousarray(requests, np.uint32)
        v_out = v_in.copy()
        m_out = m_in.copy()
        c_out = c_in.copy()
        R = max(len(req), 1)
        preq = np.zeros((R * 32, 6), self.real)
        vox_of = np.zeros(R * 32, np.uint32)
        cnt = ORcounters()
        t0 = _time.perf_counter()
        self.lib.or_lighting_mt(ctypes.byref(s), ctypes.byref(lp), req.ctypes.data_as(c_void_p), c_uint32(len(req)),
                                v_out.ctypes.data_as(c_void_p), m_out.ctypes.data_as(c_void_p),
                                c_out.ctypes.data_as(c_void_p), preq.ctypes.data_as(c_void_p),
                                vox_of.ctypes.data_as(c_void_p), ctypes.byref(cnt), c_uint32(max(1, int(threads))))
        dt = _time.perf_counter() - t0
        return v_out, m_out, c_out, preq[:len(req) * 32], vox_of[:len(req) * 32], dict(
            {f: getattr(cnt, f) for f, _ in cnt._fields_}, seconds=dt)


_cache: dict = {}


def load(kind: str = "f32") -> Oracle:
    if kind not in _cache:
        _cache[kind] = Oracle(HERE / f"liboracle_{kind}.so")
    return _cache[kind]


def golden_points(n: int) -> np.ndarray:
    """Golden-angle sphere points y = 1 - 2i/(n-1), phi = i*pi*(3 - sqrt 5), in FP64."""
    i = np.arange(n, dtype=np.float64)
    y = 1.0 - 2.0 * i / (n - 1) if n > 1 else np.ones(1)
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    phi = np.pi * (3.0 - np.sqrt(5.0)) * i
    return np.stack([np.cos(phi) * r, y, np.sin(phi) * r], axis=-1)

"""ctypes binding of oracle/_ref/libref_llvmpipe.so: the reference's GLSL shaders
run on Mesa llvmpipe (test infrastructure; see ref_llvmpipe.c). Only usable where
the reference tree and Mesa are present."""
from __future__ import annotations

import ctypes
from ctypes import Structure, c_char_p, c_double, c_float, c_size_t, c_uint32, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_ref" / "libref_llvmpipe.so"
SHADERS = Path("/root/reference/assets/shaders")


class RefScene(Structure):
    _fields_ = [("map", c_void_p), ("tiles", c_size_t), ("chunks", c_void_p), ("voxels", c_void_p),
                ("voxelCount", c_size_t), ("materials", c_void_p), ("mapSize", c_uint32 * 3),
                ("sunStrength", c_float * 3), ("ambient", c_float * 3), ("skyBot", c_float * 3),
                ("skyTop", c_float * 3)]


class RefLight(Structure):
    _fields_ = [("time", c_float), ("numDiffuseSamples", c_uint32), ("maxDiffuseSamples", c_uint32),
                ("diffuseBounceLimit", c_uint32), ("specBounceLimit", c_uint32), ("sunDir", c_float * 3),
                ("shadowSoftness", c_float), ("camPos", c_float * 3)]


class Llvmpipe:
    def __init__(self, instrument: bool = True, shader_dir: Path = SHADERS):
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `python oracle/build.py` where the reference tree is present")
        self.lib = ctypes.CDLL(str(LIB))
        self.lib.ref_error.restype = c_char_p
        self.lib.ref_renderer.restype = c_char_p
        self.lib.ref_draw.restype = c_double
        self.lib.ref_lighting.restype = c_double
        self.lib.ref_draw.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, c_void_p,
                                      c_void_p, c_void_p]
        self.lib.ref_lighting.argtypes = [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p,
                                          c_void_p]
        self.lib.ref_init.argtypes = [c_char_p, ctypes.c_int]
        r = self.lib.ref_init(str(shader_dir).encode(), int(instrument))
        if r != 0:
            raise RuntimeError(f"llvmpipe init failed ({r}): {self.lib.ref_error().decode(errors='replace')}")
        self.instrument = instrument

    def rand(self, seeds) -> np.ndarray:
        """The shader hash fract(sin(s) * 43758.5453) * 2 - 1 evaluated on llvmpipe."""
        s = np.ascontiguousarray(seeds, np.float32)
        out = np.zeros_like(s)
        self.lib.ref_rand.argtypes = [c_void_p, c_void_p, c_uint32]
        if self.lib.ref_rand(s.ctypes.data, out.ctypes.data, len(s)) != 0:
            raise RuntimeError(self.lib.ref_error().decode(errors="replace"))
        return out

    @property
    def renderer(self) -> str:
        return self.lib.ref_renderer().decode()

    def scene(self, map_size, map_, chunks, voxels, mats, sun=(0.6, 0.6, 0.6), ambient=(0.01, 0.01, 0.01),
              sky_bot=(0.71, 0.85, 0.90), sky_top=(0.0, 0.45, 0.74)):
        keep = [np.ascontiguousarray(map_, np.uint32), np.ascontiguousarray(chunks, np.uint32),
                np.ascontiguousarray(voxels, np.uint32), np.ascontiguousarray(mats)]
        s = RefScene()
        s.map, s.chunks, s.voxels, s.materials = (k.ctypes.data for k in keep)
        s.tiles = keep[0].shape[0]
        s.voxelCount = keep[2].shape[0]
        s.mapSize[:] = [int(v) for v in map_size]
        s.sunStrength[:] = sun
        s.ambient[:] = ambient
        s.skyBot[:] = sky_bot
        s.skyTop[:] = sky_top
        s._keep = keep
        return s

    def draw(self, s: RefScene, inv_view, inv_view_proj, width: int, height: int, view_mode: int = 0):
        """The draw shader with invCenteredView = invViewProj and invProjection = I, so
        its ray direction is invViewProj * (ndc, 0, 1). Returns (rgba, map_after, dbg, seconds)."""
        iv = np.ascontiguousarray(inv_view, np.float32)
        ivp = np.ascontiguousarray(inv_view_proj, np.float32)
        ident = np.eye(4, dtype=np.float32)
        rgba = np.zeros((height, width, 4), np.float32)
        mo = np.zeros((s.tiles, 3), np.uint32)
        dbg = np.zeros((height, width, 3), np.uint32)
        t = self.lib.ref_draw(ctypes.addressof(s), iv.ctypes.data, ivp.ctypes.data, ident.ctypes.data, width, height,
                              view_mode, rgba.ctypes.data, mo.ctypes.data, dbg.ctypes.data if self.instrument else None)
        return rgba, mo, dbg, t

    def lighting(self, s: RefScene, requests, *, time, num_diffuse=1, max_diffuse=1000, diffuse_bounces=5,
                 spec_bounces=2, sun_dir=(0, 1, 0), softness=10.0, cam_pos=(0, 0, 0)):
        """Returns (voxels_after, map_after, chunks_after, prequant (N,6), seconds)."""
        lp = RefLight()
        lp.time = time
        lp.numDiffuseSamples, lp.maxDiffuseSamples = num_diffuse, max_diffuse
        lp.diffuseBounceLimit, lp.specBounceLimit = diffuse_bounces, spec_bounces
        lp.sunDir[:] = [float(v) for v in sun_dir]
        lp.shadowSoftness = softness
        lp.camPos[:] = [float(v) for v in cam_pos]
        req = np.ascontiguousarray(requests, np.uint32)
        n = s.voxelCount
        vo = np.zeros((n, 4), np.uint32)
        mo = np.zeros((s.tiles, 3), np.uint32)
        co = np.zeros((s.tiles, 24), np.uint32)
        dv = np.zeros((n, 6), np.float32)
        t = self.lib.ref_lighting(ctypes.addressof(s), ctypes.addressof(lp), req.ctypes.data, len(req), vo.ctypes.data,
                                  mo.ctypes.data, co.ctypes.data, dv.ctypes.data if self.instrument else None)
        return vo, mo, co, dv, t

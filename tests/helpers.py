"""Test helpers: build reference-layout device state on the CPU with the oracle,
compare GPU exports, camera matrices without a GPU."""
from __future__ import annotations

import ctypes

import numpy as np

import oracle
from doonengine_amd import _lib
from doonengine_amd.scenes import Scene

MATERIAL_DTYPE = oracle.MATERIAL_DTYPE


def materials_array(scene: Scene) -> np.ndarray:
    m = np.zeros(256, MATERIAL_DTYPE)
    m["opacity"] = 1.0
    m["refractIndex"] = 1.0
    for i, mt in enumerate(scene.materials):
        m[i]["padding"] = (mt.emission, 0.0)
        m[i]["emissive"] = int(mt.emissive)
        m[i]["opacity"] = mt.opacity
        m[i]["specular"] = mt.specular
        m[i]["reflectType"] = mt.reflect_type
        m[i]["shininess"] = mt.shininess
    return m


def cpu_chunk(scene: Scene, cx: int, cy: int, cz: int) -> np.ndarray:
    """(8, 8, 8, 2) u32 indexed [x][y][z] like DNchunk.voxels."""
    sl = (slice(cz * 8, cz * 8 + 8), slice(cy * 8, cy * 8 + 8), slice(cx * 8, cx * 8 + 8))
    n = scene.normal_words[sl].transpose(2, 1, 0)   # (Z,Y,X) -> [x][y][z]
    a = scene.albedo_words[sl].transpose(2, 1, 0)
    return np.stack([n, a], axis=-1).astype(np.uint32)


def build_state(scene: Scene, mats: np.ndarray | None = None, node: int = 512):
    """Reference-layout state of a fully loaded scene: every non-empty chunk at
    flags 2, pool nodes of `node` voxels in map-index order."""
    o = oracle.load("f32")
    mats = materials_array(scene) if mats is None else mats
    X, Y, Z = scene.map_size
    T = X * Y * Z
    map_ = np.zeros((T, 3), np.uint32)
    chunks = np.zeros((T, 24), np.uint32)
    vox = []
    base = 0
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                t = x + X * (y + Y * z)
                cv = cpu_chunk(scene, x, y, z)
                if not ((cv[..., 0] >> 24) != 255).any():
                    continue
                ch, vv = o.chunk_to_gpu(mats, cv)
                ch[0:3] = (x, y, z)
                chunks[t] = ch
                map_[t] = (2, 0, base)
                pad = np.zeros((node, 4), np.uint32)
                pad[:len(vv)] = vv
                vox.append(pad)
                base += node
    voxels = np.concatenate(vox) if vox else np.zeros((node, 4), np.uint32)
    return map_, chunks, voxels, mats


def camera(scene: Scene, width: int | None = None, height: int | None = None, near=0.1, far=100.0):
    """(view, proj, invView, invViewProj) through the library's CPU-only helpers."""
    lib = _lib.load()
    w = width or scene.width
    h = height or scene.height
    v = _lib.DNvolume()
    v.camPos = _lib.DNvec3(*scene.cam_pos)
    v.camOrient = _lib.DNvec3(*scene.cam_orient)
    v.camFOV = 90.0
    view, proj = _lib.DNmat4(), _lib.DNmat4()
    lib.DN_set_view_projection_matrices(ctypes.byref(v), h / w, near, far, ctypes.byref(view), ctypes.byref(proj))
    iv, ivp = _lib.DNmat4(), _lib.DNmat4()
    lib.DNX_draw_matrices(view, proj, ctypes.byref(iv), ctypes.byref(ivp))
    tonp = lambda m: np.array([[m.m[c][r] for r in range(4)] for c in range(4)], np.float32)
    return view, proj, tonp(iv), tonp(ivp)

"""Python mirror of the engine API: a thin object wrapper over libdoon.so.

Method names follow the reference's C API (src/DoonEngine/voxel.h:156-371):
``sync`` is DN_sync_gpu, ``draw`` is DN_draw, ``update_lighting`` is
DN_update_lighting, and so on. Every call goes to the HIP library; there is no
Python or CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._hip import DeviceBuffer
from ._lib import DNivec3, DNmat4, DNtexture, DNuvec3, DNvec3, DNXcounters, DNXrayHit, DNXtiming

DN_READ, DN_WRITE, DN_READ_WRITE = 0, 1, 2
KIND_DRAW, KIND_LIGHTING, KIND_RAYS = 0, 1, 2

RAYHIT_DTYPE = np.dtype([("hit", "<u4"), ("voxelIndex", "<u4"), ("mapSteps", "<u4"), ("voxelSteps", "<u4"),
                         ("pos", "<f4", 3), ("normal", "<f4", 3)])
PIXELHIT_DTYPE = np.dtype([("voxelIndex", "<u4"), ("mapSteps", "<u4"), ("voxelSteps", "<u4"), ("pad", "<u4"),
                           ("face", "<f4", 3)])
MATERIAL_DTYPE = np.dtype([("padding", "<f4", 2), ("emissive", "<u4"), ("opacity", "<f4"), ("refractIndex", "<f4"),
                           ("specular", "<f4"), ("reflectType", "<u4"), ("shininess", "<u4")])


def mat_to_np(m: DNmat4) -> np.ndarray:
    """Column-major m[col][row] as a (4, 4) array indexed [col, row]."""
    return np.array([[m.m[c][r] for r in range(4)] for c in range(4)], dtype=np.float32)


def np_to_mat(a: np.ndarray) -> DNmat4:
    m = DNmat4()
    for c in range(4):
        for r in range(4):
            m.m[c][r] = float(a[c, r])
    return m


class Volume:
    """One DNvolume."""

    def __init__(self, map_size, min_chunks: int = 4096, _path: str | None = None):
        self.lib = _lib.load()
        if not self.lib.DN_init():
            raise RuntimeError("DN_init failed: no usable HIP device")
        if _path is not None:
            self.ptr = self.lib.DN_load_volume(str(_path).encode(), int(min_chunks))
            if not self.ptr:
                raise RuntimeError(f"DN_load_volume failed for {_path}")
        else:
            self.ptr = self.lib.DN_create_volume(DNuvec3(*[int(v) for v in map_size]), int(min_chunks))
            if not self.ptr:
                raise RuntimeError("DN_create_volume failed")
        self.v = self.ptr.contents
        self.map_size = (int(self.v.mapSize.x), int(self.v.mapSize.y), int(self.v.mapSize.z))
        self.tiles = int(np.prod(self.map_size))

    @classmethod
    def load(cls, path, min_chunks: int = 128) -> "Volume":
        """DN_load_volume: a volume read from a .voxvol file."""
        return cls(None, min_chunks, _path=path)

    def save(self, path) -> None:
        if not self.lib.DN_save_volume(str(path).encode(), self.ptr):
            raise RuntimeError(f"DN_save_volume failed for {path}")

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.lib.DN_delete_volume(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- materials and parameters ----------------------------------------
    @property
    def materials(self) -> np.ndarray:
        """Structured view (256,) of vol->materials (writes go straight to the volume)."""
        buf = (ctypes.c_char * (256 * 32)).from_address(ctypes.addressof(self.v.materials.contents))
        return np.frombuffer(buf, dtype=MATERIAL_DTYPE)

    def set_material(self, i: int, *, emissive=False, opacity=1.0, refract=1.0, specular=0.0, reflect_type=0,
                     shininess=0, emission=0.0) -> None:
        m = self.materials
        m[i]["padding"] = (emission, 0.0)
        m[i]["emissive"] = int(bool(emissive))
        m[i]["opacity"] = opacity
        m[i]["refractIndex"] = refract
        m[i]["specular"] = specular
        m[i]["reflectType"] = reflect_type
        m[i]["shininess"] = shininess

    def set_vec(self, name: str, xyz) -> None:
        setattr(self.v, name, DNvec3(*[float(a) for a in xyz]))

    # -- edits ------------------------------------------------------------
    def set_dense(self, normal_words: np.ndarray, albedo_words: np.ndarray, origin=(0, 0, 0)) -> None:
        """Sets a dense (Z, Y, X) box of voxels (DNX_set_voxels_dense)."""
        n = np.ascontiguousarray(normal_words, dtype=np.uint32)
        a = np.ascontiguousarray(albedo_words, dtype=np.uint32)
        assert n.shape == a.shape and n.ndim == 3
        z, y, x = n.shape
        self.lib.DNX_set_voxels_dense(self.ptr, DNivec3(*origin), DNuvec3(x, y, z), n.ctypes.data, a.ctypes.data)

    def get_dense(self, size, origin=(0, 0, 0)):
        """Reads a dense (Z, Y, X) box of voxel words (DNX_get_voxels_dense): (normal, albedo)."""
        x, y, z = (int(v) for v in size)
        n = np.empty((z, y, x), np.uint32)
        a = np.empty((z, y, x), np.uint32)
        self.lib.DNX_get_voxels_dense(self.ptr, DNivec3(*origin), DNuvec3(x, y, z), n.ctypes.data, a.ctypes.data)
        return n, a

    def chunk_exists(self, map_pos) -> bool:
        return bool(self.lib.DN_does_chunk_exist(self.ptr, DNivec3(*map_pos)))

    def set_voxel(self, map_pos, chunk_pos, material: int, normal, albedo) -> None:
        vox = _lib.DNvoxel(material, DNvec3(*normal), _lib.DNcolor(*albedo))
        self.lib.DN_set_voxel(self.ptr, DNivec3(*map_pos), DNivec3(*chunk_pos), vox)

    def remove_voxel(self, map_pos, chunk_pos) -> None:
        self.lib.DN_remove_voxel(self.ptr, DNivec3(*map_pos), DNivec3(*chunk_pos))

    def shape_sphere(self, material: int, albedo, center, radius: float, flip_normals=False,
                     orient=(0.0, 0.0, 0.0, 1.0), func=None, user_data=None) -> None:
        """DN_shape_sphere; `func` is None or the address of a C VoxelTransformFunc
        (voxelShapes.h:28), `user_data` None or an address handed to it."""
        vox = _lib.DNvoxel(material, DNvec3(0, 0, 0), _lib.DNcolor(*albedo))
        fp = None if func is None else ctypes.c_void_p(func if isinstance(func, int) else
                                                       ctypes.cast(func, ctypes.c_void_p).value)
        ud = None if user_data is None else ctypes.c_void_p(user_data)
        self.lib.DN_shape_sphere(self.ptr, vox, bool(flip_normals), DNvec3(*center), float(radius),
                                 _lib.DNquaternion(*orient), fp, ud)

    @staticmethod
    def sphere_edits(edits):
        """Packs (material, albedo, center, radius, flip_normals) tuples into a
        DNXsphereEdit array for shape_spheres."""
        arr = (_lib.DNXsphereEdit * max(1, len(edits)))()
        for i, (material, albedo, center, radius, flip) in enumerate(edits):
            arr[i].voxel = _lib.DNvoxel(material, DNvec3(0, 0, 0), _lib.DNcolor(*albedo))
            arr[i].center = DNvec3(*center)
            arr[i].radius = float(radius)
            arr[i].flipNormals = 1 if flip else 0
        return arr, len(edits)

    def shape_spheres(self, edits) -> None:
        """Applies sphere edits in order on the GPU (DNX_shape_spheres); `edits` is a
        sequence of (material, albedo, center, radius, flip_normals) or a packed
        (array, count) pair from sphere_edits."""
        arr, n = edits if isinstance(edits, tuple) else self.sphere_edits(edits)
        if not self.lib.DNX_shape_spheres(self.ptr, arr, n):
            raise RuntimeError("DNX_shape_spheres failed")

    def host_state(self):
        """Copies of the CPU map handles and chunk array (DNvolume.map / .chunks) as
        numpy structured arrays."""
        v = self.ptr.contents
        tiles = v.mapSize.x * v.mapSize.y * v.mapSize.z
        hsz, csz = ctypes.sizeof(_lib.DNchunkHandle), ctypes.sizeof(_lib.DNchunk)
        mp = np.frombuffer(ctypes.string_at(ctypes.addressof(v.map.contents), tiles * hsz),
                           dtype=np.dtype({"names": ["flag", "chunkIndex"], "formats": ["u1", "<u4"],
                                           "offsets": [0, 4], "itemsize": hsz}))
        ch = np.frombuffer(ctypes.string_at(ctypes.addressof(v.chunks.contents), v.chunkCap * csz),
                           dtype=np.dtype({"names": ["pos", "updated", "numVoxels", "numVoxelsGpu", "voxels"],
                                           "formats": [("<i4", 3), "u1", "<u4", "<u4", ("<u4", (512, 2))],
                                           "offsets": [0, 12, 16, 20, 24], "itemsize": csz}))
        return mp.copy(), ch.copy(), int(v.nextChunk)

    def set_sky_cubemap(self, faces) -> None:
        """DNX_set_sky_cubemap: faces (6, n, n, 3) uint8 in GL order (+X, -X, +Y, -Y,
        +Z, -Z), row 0 first; None removes the cubemap. Sets useCubemap."""
        if faces is None:
            self.lib.DNX_set_sky_cubemap(self.ptr, None, 0)
            self.cube_faces = None
            self.v.useCubemap = False
            return
        f = np.ascontiguousarray(faces, np.uint8)
        assert f.ndim == 4 and f.shape[0] == 6 and f.shape[1] == f.shape[2] and f.shape[3] == 3
        if not self.lib.DNX_set_sky_cubemap(self.ptr, f.ctypes.data, f.shape[1]):
            raise RuntimeError("DNX_set_sky_cubemap failed")
        self.cube_faces = f
        self.v.useCubemap = True

    def pick(self, direction, origin, max_steps: int):
        """DN_step_map (the CPU pick ray): (hit, voxel position, face normal, DNvoxel)."""
        pos, nrm, vox = DNivec3(), DNivec3(), _lib.DNvoxel()
        hit = self.lib.DN_step_map(self.ptr, DNvec3(*direction), DNvec3(*origin), int(max_steps), ctypes.byref(pos),
                                   ctypes.byref(vox), ctypes.byref(nrm))
        return bool(hit), (pos.x, pos.y, pos.z), (nrm.x, nrm.y, nrm.z), vox

    def request_all(self) -> None:
        self.lib.DNX_request_all_chunks(self.ptr)

    def set_pool_limit(self, max_voxels: int) -> bool:
        return bool(self.lib.DNX_set_voxel_pool_limit(self.ptr, int(max_voxels)))

    def residency_stats(self, reset: bool = False) -> dict:
        e, d = ctypes.c_uint64(), ctypes.c_uint64()
        self.lib.DNX_get_residency_stats(self.ptr, ctypes.byref(e), ctypes.byref(d), bool(reset))
        return {"evictions": e.value, "deferrals": d.value}

    def mark_visible(self, max_chunks: int) -> int:
        return int(self.lib.DNX_mark_visible(self.ptr, int(max_chunks)))

    # -- pipeline ---------------------------------------------------------
    def sync(self, op: int = DN_READ_WRITE, split: int = 1) -> None:
        self.lib.DN_sync_gpu(self.ptr, op, split)

    def view_projection(self, aspect: float, near: float = 0.1, far: float = 100.0):
        view, proj = DNmat4(), DNmat4()
        self.lib.DN_set_view_projection_matrices(self.ptr, aspect, near, far, ctypes.byref(view), ctypes.byref(proj))
        return view, proj

    def draw(self, out: DeviceBuffer, width: int, height: int, view: DNmat4, proj: DNmat4) -> None:
        assert out.nbytes >= width * height * 16
        self.lib.DN_draw(self.ptr, DNtexture(out.ptr, width, height), view, proj, -1, -1)

    def update_lighting(self, num_diffuse: int = 1, max_diffuse: int = 1000, time: float = 123.456) -> None:
        self.lib.DN_update_lighting(self.ptr, num_diffuse, max_diffuse, time)

    def synchronize(self) -> None:
        if not self.lib.DNX_synchronize(self.ptr):
            raise RuntimeError("device error")

    def cast_rays(self, origins: np.ndarray, dirs: np.ndarray, ignore_first: bool = False) -> np.ndarray:
        o = np.ascontiguousarray(origins, dtype=np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = o.shape[0]
        bo, bd = DeviceBuffer.from_array(o), DeviceBuffer.from_array(d)
        bout = DeviceBuffer(n * RAYHIT_DTYPE.itemsize)
        self.lib.DNX_cast_rays(self.ptr, n, bo.ptr, bd.ptr, bool(ignore_first), bout.ptr)
        self.synchronize()
        return bout.download(RAYHIT_DTYPE, (n,))

    # -- inspection -------------------------------------------------------
    def export_state(self):
        """Device state in the reference layouts: map (T,3), chunks (T,24), voxels (cap,4), all uint32."""
        T = self.tiles
        cap = int(self.v.voxelCap)
        m = np.zeros((T, 3), np.uint32)
        c = np.zeros((T, 24), np.uint32)
        vx = np.zeros((cap, 4), np.uint32)
        self.lib.DNX_export_gpu_state(self.ptr, m.ctypes.data, c.ctypes.data, vx.ctypes.data, vx.size)
        return m, c, vx

    def fetch_requests(self) -> np.ndarray:
        n = self.lib.DNX_fetch_lighting_requests(self.ptr)
        return np.ctypeslib.as_array(self.v.lightingRequests, shape=(max(n, 1),))[:n].copy()

    def counters(self, kind: int, reset: bool = False) -> dict:
        c = DNXcounters()
        self.lib.DNX_get_counters(self.ptr, kind, ctypes.byref(c), reset)
        return {f: getattr(c, f) for f, _ in c._fields_}

    def enable_timing(self, on: bool = True) -> None:
        self.lib.DNX_enable_timing(self.ptr, on)

    def timing(self, reset: bool = False) -> dict:
        t = DNXtiming()
        self.lib.DNX_get_timing(self.ptr, ctypes.byref(t), reset)
        return {f: getattr(t, f) for f, _ in t._fields_}

    def ray_points(self, which: int) -> np.ndarray:
        out = np.zeros((64, 3), np.float32)
        n = self.lib.DNX_get_ray_points(self.ptr, which, out.ctypes.data, 64)
        return out[:n].copy()


class Batch:
    """DNXbatch: the lighting pass of many volumes in one persistent, graph-replayed launch."""

    def __init__(self, volumes):
        self.lib = _lib.load()
        self.volumes = list(volumes)
        arr = (ctypes.POINTER(_lib.DNvolume) * len(self.volumes))(*[v.ptr for v in self.volumes])
        self.ptr = self.lib.DNX_batch_create(arr, len(self.volumes))
        if not self.ptr:
            raise RuntimeError("DNX_batch_create failed")

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.lib.DNX_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def config(self) -> dict:
        c = _lib.DNXbatchConfig()
        self.lib.DNX_batch_get_config(self.ptr, ctypes.byref(c))
        return {f: getattr(c, f) for f, _ in c._fields_}

    def set_config(self, requests_per_unit=8, block_threads=256, blocks_per_cu=0, xcd_queues=1) -> None:
        c = _lib.DNXbatchConfig(requests_per_unit, block_threads, blocks_per_cu, xcd_queues)
        if not self.lib.DNX_batch_set_config(self.ptr, ctypes.byref(c)):
            raise ValueError("invalid batch configuration")

    def update_lighting(self, num_diffuse: int = 1, max_diffuse: int = 1000, time: float = 123.456) -> None:
        if not self.lib.DNX_batch_update_lighting(self.ptr, num_diffuse, max_diffuse, time):
            raise RuntimeError("DNX_batch_update_lighting failed")

    def autotune(self, num_diffuse: int = 1, max_diffuse: int = 1000, time: float = 123.456, repeats: int = 3):
        c, ms = _lib.DNXbatchConfig(), ctypes.c_float()
        if not self.lib.DNX_batch_autotune(self.ptr, num_diffuse, max_diffuse, time, repeats, ctypes.byref(c),
                                           ctypes.byref(ms)):
            raise RuntimeError("DNX_batch_autotune failed")
        return {f: getattr(c, f) for f, _ in c._fields_}, ms.value

    def timing(self, reset: bool = False):
        n = ctypes.c_uint32()
        ms = self.lib.DNX_batch_timing(self.ptr, ctypes.byref(n), reset)
        return ms, n.value


RC_NPARAM = 12288
RC_OFFSETS = ((0, 64, 48, 38), (3072, 64, 64, 64), (7168, 64, 64, 64), (11264, 16, 64, 3))  # off, rows, ld, used


def rc_pack_weights(mats) -> np.ndarray:
    """[W0 (64x38), W1, W2 (64x64), W3 (3x64)] -> the padded 12288-float layout."""
    out = np.zeros(RC_NPARAM, np.float32)
    for (off, rows, ld, _), m in zip(RC_OFFSETS, mats):
        blk = np.zeros((rows, ld), np.float32)
        blk[:m.shape[0], :m.shape[1]] = m
        out[off:off + rows * ld] = blk.reshape(-1)
    return out


def rc_unpack_weights(w: np.ndarray):
    mats = []
    for off, rows, ld, used in RC_OFFSETS:
        blk = np.asarray(w[off:off + rows * ld]).reshape(rows, ld)
        mats.append(blk[:3, :] if used == 3 else blk[:, :used])
    return mats


class RadianceCache:
    """DNXradianceCache of one volume."""

    def __init__(self, vol: Volume, **cfg):
        self.lib = _lib.load()
        self.vol = vol
        c = _lib.DNXrcConfig()
        self.lib.DNX_rc_default_config(ctypes.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.ptr = self.lib.DNX_rc_create(vol.ptr, ctypes.byref(c))
        if not self.ptr:
            raise RuntimeError("DNX_rc_create failed")
        self.config = {f: getattr(c, f) for f, _ in c._fields_}
        self.table_floats = 16 * 2 * (1 << c.log2Table)

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.lib.DNX_rc_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train_frame(self, time: float, steps: int = 1) -> None:
        if not self.lib.DNX_rc_train_frame(self.ptr, time, steps):
            raise RuntimeError("DNX_rc_train_frame failed")

    def predict_requests(self, out: DeviceBuffer, request_capacity: int) -> None:
        if not self.lib.DNX_rc_predict_requests(self.ptr, out.ptr, request_capacity):
            raise RuntimeError("DNX_rc_predict_requests failed")

    def reference(self, acc: DeviceBuffer, request_capacity: int, samples_per_voxel: int, seed: int = 7) -> None:
        """acc (4 doubles per slot: rgb sums, count) += independent samples of every voxel."""
        if not self.lib.DNX_rc_reference(self.ptr, acc.ptr, request_capacity, samples_per_voxel, seed):
            raise RuntimeError("DNX_rc_reference failed")

    def train_samples(self, samples: DeviceBuffer, n: int, step: bool) -> None:
        if not self.lib.DNX_rc_train_samples(self.ptr, samples.ptr, n, step):
            raise RuntimeError("DNX_rc_train_samples failed")

    def predict_samples(self, samples: DeviceBuffer, n: int, out: DeviceBuffer) -> None:
        if not self.lib.DNX_rc_predict_samples(self.ptr, samples.ptr, n, out.ptr):
            raise RuntimeError("DNX_rc_predict_samples failed")

    def predict_samples_noalloc(self, samples: DeviceBuffer, n: int, out: DeviceBuffer) -> bool:
        """The allocation-free variant (scratch rows in a fixed per-device array);
        False when n exceeds DNX_RC_NOALLOC_MAX_SAMPLES or is not a multiple of 16."""
        return bool(self.lib.DNX_rc_predict_samples_noalloc(self.ptr, samples.ptr, n, out.ptr))

    def set_params(self, weights: np.ndarray, table: np.ndarray) -> None:
        w = np.ascontiguousarray(weights, np.float32)
        t = np.ascontiguousarray(table, np.float32)
        assert w.size == RC_NPARAM and t.size == self.table_floats
        if not self.lib.DNX_rc_set_params(self.ptr, w.ctypes.data, t.ctypes.data):
            raise RuntimeError("DNX_rc_set_params failed")

    def get_params(self):
        """(weights, table, weight gradients, table gradients) as float32 arrays."""
        w, g = np.zeros(RC_NPARAM, np.float32), np.zeros(RC_NPARAM, np.float32)
        t, tg = np.zeros(self.table_floats, np.float32), np.zeros(self.table_floats, np.float32)
        if not self.lib.DNX_rc_get_params(self.ptr, w.ctypes.data, t.ctypes.data, g.ctypes.data, tg.ctypes.data):
            raise RuntimeError("DNX_rc_get_params failed")
        return w, t, g, tg

    def last_samples(self) -> np.ndarray:
        n = ctypes.c_uint32()
        p = self.lib.DNX_rc_last_samples(self.ptr, ctypes.byref(n))
        out = np.zeros((n.value, 16), np.float32)
        if n.value:
            from ._hip import hip, check
            self.vol.synchronize()
            check(hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(p), out.nbytes, 2), "samples")
        return out


def draw_matrices(view: DNmat4, proj: DNmat4):
    """(invView, invCenteredView * invProjection) as the draw kernel receives them."""
    lib = _lib.load()
    iv, ivp = DNmat4(), DNmat4()
    lib.DNX_draw_matrices(view, proj, ctypes.byref(iv), ctypes.byref(ivp))
    return iv, ivp

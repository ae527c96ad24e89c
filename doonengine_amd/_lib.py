"""ctypes bindings of libdoon.so (include/doon/doon.h).

The structures below mirror the C-ABI; ``check_layout()`` compares their sizes
with the library's own ``DNX_struct_size``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_bool, c_char_p, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libdoon.so"


class DNivec3(Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32)]


class DNuvec3(Structure):
    _fields_ = [("x", c_uint32), ("y", c_uint32), ("z", c_uint32)]


class DNvec3(Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


class DNvec2(Structure):
    _fields_ = [("x", c_float), ("y", c_float)]


class DNmat4(Structure):
    _fields_ = [("m", (c_float * 4) * 4)]


class DNquaternion(Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float), ("w", c_float)]


class DNcolor(Structure):
    _fields_ = [("r", ctypes.c_uint8), ("g", ctypes.c_uint8), ("b", ctypes.c_uint8)]


class DNvoxel(Structure):
    _fields_ = [("material", ctypes.c_uint8), ("normal", DNvec3), ("albedo", DNcolor)]


class DNcompressedVoxel(Structure):
    _fields_ = [("normal", c_uint32), ("albedo", c_uint32)]


class DNchunk(Structure):
    _fields_ = [("pos", DNivec3), ("updated", c_bool), ("numVoxels", c_uint32), ("numVoxelsGpu", c_uint32),
                ("voxels", ((DNcompressedVoxel * 8) * 8) * 8)]


class DNchunkHandle(Structure):
    _fields_ = [("flag", ctypes.c_uint8), ("chunkIndex", c_uint32)]


class DNvoxelNode(Structure):
    _fields_ = [("size", c_uint32), ("startPos", c_size_t), ("chunkPos", DNivec3)]


class DNmaterial(Structure):
    _fields_ = [("padding", DNvec2), ("emissive", c_uint32), ("opacity", c_float), ("refractIndex", c_float),
                ("specular", c_float), ("reflectType", c_uint32), ("shininess", c_uint32)]


class DNtexture(Structure):
    _fields_ = [("data", c_void_p), ("width", c_uint32), ("height", c_uint32)]


class DNvolume(Structure):
    _fields_ = [
        ("gpuMap", c_void_p), ("gpuChunks", c_void_p), ("gpuVoxels", c_void_p),
        ("mapSize", DNuvec3),
        ("chunkCap", c_size_t), ("nextChunk", c_size_t), ("voxelCap", c_size_t), ("numVoxelNodes", c_size_t),
        ("numLightingRequests", c_size_t), ("lightingRequestCap", c_size_t),
        ("map", POINTER(DNchunkHandle)), ("chunks", POINTER(DNchunk)), ("materials", POINTER(DNmaterial)),
        ("lightingRequests", POINTER(c_uint32)), ("gpuVoxelLayout", POINTER(DNvoxelNode)),
        ("camPos", DNvec3), ("camOrient", DNvec3), ("camFOV", c_float), ("camViewMode", c_uint32),
        ("sunDir", DNvec3), ("sunStrength", DNvec3), ("ambientLightStrength", DNvec3),
        ("diffuseBounceLimit", c_uint32), ("specBounceLimit", c_uint32), ("shadowSoftness", c_float),
        ("useCubemap", c_bool), ("cubemapTex", c_uint32), ("skyGradientBot", DNvec3), ("skyGradientTop", DNvec3),
        ("frameNum", c_uint32), ("lastTime", c_float),
        ("numSpecularRays", c_uint32), ("numDirectRays", c_uint32), ("gpu", c_void_p),
    ]


class DNXcounters(Structure):
    _fields_ = [("rays", c_uint64), ("mapSteps", c_uint64), ("voxelSteps", c_uint64), ("hits", c_uint64),
                ("litVoxels", c_uint64), ("pixels", c_uint64)]


class DNXtiming(Structure):
    _fields_ = [("drawMs", ctypes.c_double), ("lightingMs", ctypes.c_double), ("visibleMs", ctypes.c_double),
                ("uploadMs", ctypes.c_double), ("epilogueMs", ctypes.c_double), ("raysMs", ctypes.c_double),
                ("drawLaunches", c_uint32), ("lightingLaunches", c_uint32), ("visibleLaunches", c_uint32),
                ("uploadLaunches", c_uint32), ("epilogueLaunches", c_uint32), ("raysLaunches", c_uint32)]


class DNXrayHit(Structure):
    _fields_ = [("hit", c_uint32), ("voxelIndex", c_uint32), ("mapSteps", c_uint32), ("voxelSteps", c_uint32),
                ("pos", c_float * 3), ("normal", c_float * 3)]


class DNXpixelHit(Structure):
    _fields_ = [("voxelIndex", c_uint32), ("mapSteps", c_uint32), ("voxelSteps", c_uint32), ("pad", c_uint32),
                ("face", c_float * 3)]


class DNXbatchConfig(Structure):
    _fields_ = [("requestsPerUnit", c_uint32), ("blockThreads", c_uint32), ("blocksPerCU", c_uint32),
                ("xcdQueues", c_uint32)]


class DNXsphereEdit(Structure):
    _fields_ = [("voxel", DNvoxel), ("center", DNvec3), ("radius", c_float), ("flipNormals", c_uint32)]


class DNXrcConfig(Structure):
    _fields_ = [("log2Table", c_uint32), ("nMin", c_uint32), ("nMax", c_uint32), ("extent", c_float),
                ("lr", c_float), ("beta1", c_float), ("beta2", c_float), ("eps", c_float),
                ("samplesPerFrame", c_uint32), ("seed", c_uint32), ("emaDecay", c_float),
                ("lrDecaySteps", c_float), ("lrDecayPower", c_float), ("emaGrowth", c_float)]


MSG_CB = ctypes.CFUNCTYPE(None, c_int, c_int, c_char_p)

# name: (restype, argtypes)
SIGNATURES = {
    "DN_init": (c_bool, []),
    "DN_quit": (None, []),
    "DN_create_volume": (POINTER(DNvolume), [DNuvec3, ctypes.c_uint]),
    "DN_delete_volume": (None, [POINTER(DNvolume)]),
    "DN_set_view_projection_matrices": (None, [POINTER(DNvolume), c_float, c_float, c_float, POINTER(DNmat4),
                                               POINTER(DNmat4)]),
    "DN_draw": (None, [POINTER(DNvolume), DNtexture, DNmat4, DNmat4, c_int, c_int]),
    "DN_update_lighting": (None, [POINTER(DNvolume), c_int, c_int, c_float]),
    "DN_add_chunk": (c_int, [POINTER(DNvolume), DNivec3]),
    "DN_remove_chunk": (None, [POINTER(DNvolume), DNivec3]),
    "DN_sync_gpu": (None, [POINTER(DNvolume), c_int, c_int]),
    "DN_set_map_size": (c_bool, [POINTER(DNvolume), DNuvec3]),
    "DN_set_max_chunks": (c_bool, [POINTER(DNvolume), c_size_t]),
    "DN_set_max_voxels_gpu": (c_bool, [POINTER(DNvolume), c_size_t]),
    "DN_set_max_lighting_requests": (c_bool, [POINTER(DNvolume), c_size_t]),
    "DN_in_map_bounds": (c_bool, [POINTER(DNvolume), DNivec3]),
    "DN_in_chunk_bounds": (c_bool, [DNivec3]),
    "DN_get_voxel": (DNvoxel, [POINTER(DNvolume), DNivec3, DNivec3]),
    "DN_get_compressed_voxel": (DNcompressedVoxel, [POINTER(DNvolume), DNivec3, DNivec3]),
    "DN_set_voxel": (None, [POINTER(DNvolume), DNivec3, DNivec3, DNvoxel]),
    "DN_set_compressed_voxel": (None, [POINTER(DNvolume), DNivec3, DNivec3, DNcompressedVoxel]),
    "DN_remove_voxel": (None, [POINTER(DNvolume), DNivec3, DNivec3]),
    "DN_does_chunk_exist": (c_bool, [POINTER(DNvolume), DNivec3]),
    "DN_does_voxel_exist": (c_bool, [POINTER(DNvolume), DNivec3, DNivec3]),
    "DN_step_map": (c_bool, [POINTER(DNvolume), DNvec3, DNvec3, c_int, POINTER(DNivec3), POINTER(DNvoxel),
                             POINTER(DNivec3)]),
    "DN_separate_position": (None, [DNivec3, POINTER(DNivec3), POINTER(DNivec3)]),
    "DN_cam_dir": (DNvec3, [DNvec3]),
    "DN_compress_voxel": (DNcompressedVoxel, [DNvoxel]),
    "DN_decompress_voxel": (DNvoxel, [DNcompressedVoxel]),
    "DN_shape_sphere": (None, [POINTER(DNvolume), DNvoxel, c_bool, DNvec3, c_float, DNquaternion, c_void_p, c_void_p]),
    "DNX_struct_size": (c_size_t, [c_int]),
    "DNX_set_message_callback": (None, [MSG_CB]),
    "DNX_error_count": (c_uint32, []),
    "DNX_set_voxels_dense": (None, [POINTER(DNvolume), DNivec3, DNuvec3, c_void_p, c_void_p]),
    "DNX_get_voxels_dense": (None, [POINTER(DNvolume), DNivec3, DNuvec3, c_void_p, c_void_p]),
    "DNX_shape_spheres": (c_bool, [POINTER(DNvolume), POINTER(DNXsphereEdit), c_uint32]),
    "DNX_set_sky_cubemap": (c_bool, [POINTER(DNvolume), c_void_p, c_uint32]),
    "DN_load_volume": (POINTER(DNvolume), [ctypes.c_char_p, c_uint32]),
    "DN_save_volume": (c_bool, [ctypes.c_char_p, POINTER(DNvolume)]),
    "DNX_request_all_chunks": (None, [POINTER(DNvolume)]),
    "DNX_set_voxel_pool_limit": (c_bool, [POINTER(DNvolume), c_size_t]),
    "DNX_get_residency_stats": (None, [POINTER(DNvolume), POINTER(c_uint64), POINTER(c_uint64), c_bool]),
    "DNX_mark_visible": (c_uint32, [POINTER(DNvolume), c_uint32]),
    "DNX_get_stream": (c_void_p, [POINTER(DNvolume)]),
    "DNX_synchronize": (c_bool, [POINTER(DNvolume)]),
    "DNX_fetch_lighting_requests": (c_size_t, [POINTER(DNvolume)]),
    "DNX_export_gpu_state": (c_size_t, [POINTER(DNvolume), c_void_p, c_void_p, c_void_p, c_size_t]),
    "DNX_get_ray_points": (c_uint32, [POINTER(DNvolume), c_int, c_void_p, c_uint32]),
    "DNX_get_counters": (None, [POINTER(DNvolume), c_int, POINTER(DNXcounters), c_bool]),
    "DNX_enable_timing": (None, [POINTER(DNvolume), c_bool]),
    "DNX_get_timing": (None, [POINTER(DNvolume), POINTER(DNXtiming), c_bool]),
    "DNX_cast_rays": (None, [POINTER(DNvolume), c_uint32, c_void_p, c_void_p, c_bool, c_void_p]),
    "DNX_set_draw_hit_output": (None, [POINTER(DNvolume), c_void_p, c_size_t]),
    "DNX_set_lighting_debug_output": (None, [POINTER(DNvolume), c_void_p, c_void_p, c_size_t]),
    "DNX_set_lighting_sample_output": (None, [POINTER(DNvolume), c_void_p, c_size_t]),
    "DNX_get_upload_stats": (None, [POINTER(DNvolume), POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64), c_bool]),
    "DNX_draw_matrices": (None, [DNmat4, DNmat4, POINTER(DNmat4), POINTER(DNmat4)]),
    "DNX_batch_create": (c_void_p, [POINTER(POINTER(DNvolume)), c_uint32]),
    "DNX_batch_destroy": (None, [c_void_p]),
    "DNX_batch_set_config": (c_bool, [c_void_p, POINTER(DNXbatchConfig)]),
    "DNX_batch_get_config": (None, [c_void_p, POINTER(DNXbatchConfig)]),
    "DNX_batch_update_lighting": (c_bool, [c_void_p, c_int, c_int, c_float]),
    "DNX_batch_autotune": (c_bool, [c_void_p, c_int, c_int, c_float, c_int, POINTER(DNXbatchConfig),
                                    POINTER(c_float)]),
    "DNX_batch_timing": (ctypes.c_double, [c_void_p, POINTER(c_uint32), c_bool]),
    "DNX_rc_default_config": (None, [POINTER(DNXrcConfig)]),
    "DNX_rc_create": (c_void_p, [POINTER(DNvolume), POINTER(DNXrcConfig)]),
    "DNX_rc_destroy": (None, [c_void_p]),
    "DNX_rc_train_frame": (c_bool, [c_void_p, c_float, c_int]),
    "DNX_rc_predict_requests": (c_bool, [c_void_p, c_void_p, c_size_t]),
    "DNX_rc_train_samples": (c_bool, [c_void_p, c_void_p, c_uint32, c_bool]),
    "DNX_rc_predict_samples": (c_bool, [c_void_p, c_void_p, c_uint32, c_void_p]),
    "DNX_rc_predict_samples_noalloc": (c_bool, [c_void_p, c_void_p, c_uint32, c_void_p]),
    "DNX_rc_set_params": (c_bool, [c_void_p, c_void_p, c_void_p]),
    "DNX_rc_get_params": (c_bool, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "DNX_rc_last_samples": (c_void_p, [c_void_p, POINTER(c_uint32)]),
    "DNX_rc_reference": (c_bool, [c_void_p, c_void_p, c_size_t, c_uint32, c_uint32]),
    "DNX_accumulate_f64": (c_bool, [POINTER(DNvolume), c_void_p, c_void_p, c_size_t]),
}

STRUCT_IDS = {0: DNvolume, 1: DNchunk, 2: DNmaterial, 3: DNchunkHandle, 4: DNvoxelNode, 5: DNXcounters,
              6: DNXtiming, 7: DNXrayHit, 8: DNXpixelHit, 9: DNvoxel, 12: DNXbatchConfig, 13: DNXrcConfig,
              14: DNXsphereEdit}

_lib = None


def load(path: Path | str | None = None) -> ctypes.CDLL:
    """Loads libdoon.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    # DOON_LIB selects another build of the library (e.g. a variant built with
    # different kernel options) for side-by-side measurements.
    p = Path(path) if path else Path(os.environ.get("DOON_LIB", str(LIB_PATH)))
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -m doonengine_amd.build`")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check_layout(lib: ctypes.CDLL) -> None:
    for k, st in STRUCT_IDS.items():
        n = lib.DNX_struct_size(k)
        if n != ctypes.sizeof(st):
            raise RuntimeError(f"layout mismatch for {st.__name__}: C {n}, ctypes {ctypes.sizeof(st)}")
    if lib.DNX_struct_size(10) != DNvolume.numSpecularRays.offset or lib.DNX_struct_size(11) != DNvolume.camPos.offset:
        raise RuntimeError("layout mismatch inside DNvolume")

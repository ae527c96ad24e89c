"""Seeded synthetic scenes for the voxel pipeline (numpy).

Terrain: an fBm value-noise heightfield (4 octaves, lacunarity 2, gain 0.5)
whose base height is bisected to a target occupancy; optional axis-aligned
boxes. Every voxel carries a material id, an 8-bit albedo and a normal encoded
as DN_compress_voxel does (voxel.c:1289-1301). All randomness comes from an
integer hash of coordinates and a seed, so scenes are identical everywhere.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

EMPTY = 255


def hash_u32(*keys, seed: int = 0) -> np.ndarray:
    """Avalanche hash of integer arrays (murmur3 finaliser over a running mix)."""
    h0 = np.uint64((seed * 0x9E3779B1 + 0x85EBCA77) & 0xFFFFFFFF)
    h = np.full(np.broadcast_shapes(*[np.shape(k) for k in keys]), h0, dtype=np.uint64)
    for k in keys:
        k = np.asarray(k).astype(np.int64).astype(np.uint64) & np.uint64(0xFFFFFFFF)
        h = (h ^ k) * np.uint64(0xCC9E2D51) & np.uint64(0xFFFFFFFF)
        h = ((h << np.uint64(15)) | (h >> np.uint64(17))) & np.uint64(0xFFFFFFFF)
        h = (h * np.uint64(0x1B873593)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    return h.astype(np.uint32)


def uniform01(*keys, seed: int = 0) -> np.ndarray:
    return hash_u32(*keys, seed=seed).astype(np.float64) / 4294967296.0


def value_noise(x: np.ndarray, z: np.ndarray, octave: int, seed: int) -> np.ndarray:
    ix, iz = np.floor(x).astype(np.int64), np.floor(z).astype(np.int64)
    fx, fz = x - ix, z - iz
    sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
    v00 = uniform01(ix, iz, octave, seed=seed)
    v10 = uniform01(ix + 1, iz, octave, seed=seed)
    v01 = uniform01(ix, iz + 1, octave, seed=seed)
    v11 = uniform01(ix + 1, iz + 1, octave, seed=seed)
    a = v00 + (v10 - v00) * sx
    b = v01 + (v11 - v01) * sx
    return a + (b - a) * sz


def fbm(x, z, seed: int, octaves: int = 4, lacunarity: float = 2.0, gain: float = 0.5) -> np.ndarray:
    total, amp, norm, freq = 0.0, 1.0, 0.0, 1.0
    for o in range(octaves):
        total = total + amp * (2.0 * value_noise(x * freq, z * freq, o, seed) - 1.0)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def encode_normal(n: np.ndarray) -> np.ndarray:
    """DN_compress_voxel's normal bytes: ((int)(n*255) + 255) / 2 per component."""
    n = np.clip(n.astype(np.float32), -1.0, 1.0)
    q = np.trunc(n * np.float32(255.0)).astype(np.int32)
    return ((q + 255) // 2).astype(np.uint32)


def pack_normal_word(material: np.ndarray, n: np.ndarray) -> np.ndarray:
    b = encode_normal(n)
    return (material.astype(np.uint32) << 24) | (b[..., 0] << 16) | (b[..., 1] << 8) | b[..., 2]


@dataclass
class Material:
    albedo: tuple            # 8-bit RGB
    emissive: bool = False
    emission: float = 0.0    # > 0 overrides the reference's emission = albedo
    specular: float = 0.0
    shininess: int = 0
    reflect_type: int = 0
    opacity: float = 1.0


@dataclass
class Scene:
    name: str
    map_size: tuple                      # chunks (x, y, z)
    normal_words: np.ndarray             # (Z, Y, X) voxels
    albedo_words: np.ndarray
    materials: list
    cam_pos: tuple                       # chunks
    cam_orient: tuple                    # degrees (pitch, yaw, roll)
    sun_dir: tuple = (-1.0, 1.0, -1.0)
    width: int = 256
    height: int = 256
    params: dict = field(default_factory=dict)
    visible: int = 0                     # chunks marked visible per frame (batched workload)

    @property
    def occupancy(self) -> float:
        return float(((self.normal_words >> 24) != EMPTY).mean())

    def apply(self, vol) -> None:
        """Loads the scene into a Volume: materials, voxels, camera and sun."""
        for i, m in enumerate(self.materials):
            vol.set_material(i, emissive=m.emissive, emission=m.emission, specular=m.specular,
                             shininess=m.shininess, reflect_type=m.reflect_type, opacity=m.opacity)
        vol.set_dense(self.normal_words, self.albedo_words)
        vol.set_vec("camPos", self.cam_pos)
        vol.set_vec("camOrient", self.cam_orient)
        vol.set_vec("sunDir", self.sun_dir)
        for k, v in self.params.items():
            setattr(vol.v, k, v)


def _materials(n: int, seed: int, emissive_index: int | None, emission: float) -> list:
    mats = []
    for i in range(n):
        rgb = tuple(int(round(255 * (0.2 + 0.7 * uniform01(i, c, seed=seed)))) for c in range(3))
        mats.append(Material(albedo=rgb))
    if emissive_index is not None:
        mats[emissive_index].emissive = True
        mats[emissive_index].emission = emission
    return mats


def terrain(size_vox, occupancy: float, seed: int = 1, amplitude: float | None = None, n_boxes: int = 0,
            box_seed: int = 3, box_size=(2, 40), n_materials: int = 4, emissive_fraction: float = 0.005,
            emission: float = 4.0):
    """Dense voxel arrays (Z, Y, X) of an fBm terrain plus boxes."""
    X, Y, Z = size_vox
    xs = np.arange(X, dtype=np.float64)
    zs = np.arange(Z, dtype=np.float64)
    gx, gz = np.meshgrid(xs, zs)                     # (Z, X)
    f = fbm(gx / 64.0, gz / 64.0, seed)
    amp = amplitude if amplitude is not None else 0.35 * Y
    # bisect the base height to the target occupancy of the heightfield alone
    lo, hi = -amp, float(Y)
    for _ in range(60):
        h0 = 0.5 * (lo + hi)
        occ = np.clip(h0 + amp * f, 0, Y).mean() / Y
        if occ < occupancy:
            lo = h0
        else:
            hi = h0
    h = h0 + amp * f                                 # (Z, X) float heights
    hz, hx = np.gradient(h)
    nrm = np.stack([-hx, np.ones_like(h), -hz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)

    ys = np.arange(Y)
    solid = ys[None, :, None] < h[:, None, :]        # (Z, Y, X)
    diffuse_mats = max(1, n_materials - 1)
    q = np.quantile(h, [1 / 3, 2 / 3]) if diffuse_mats >= 3 else [np.inf, np.inf]
    band = np.where(h < q[0], 0, np.where(h < q[1], 1, 2)) % diffuse_mats
    mat = np.broadcast_to(band[:, None, :], solid.shape).astype(np.uint32).copy()
    normals = np.broadcast_to(nrm[:, None, :, :], solid.shape + (3,)).astype(np.float32).copy()

    for b in range(n_boxes):
        sx, sy, sz = (int(box_size[0] + (box_size[1] - box_size[0]) * uniform01(b, k, seed=box_seed)) for k in range(3))
        ox = int(uniform01(b, 3, seed=box_seed) * max(1, X - sx))
        oz = int(uniform01(b, 5, seed=box_seed) * max(1, Z - sz))
        col = h[oz:oz + sz, ox:ox + sx]
        oy = int(np.clip(col.mean() - sy * 0.3, 0, max(0, Y - sy)))
        sl = (slice(oz, oz + sz), slice(oy, oy + sy), slice(ox, ox + sx))
        solid[sl] = True
        mat[sl] = b % diffuse_mats
        lz, ly, lx = np.meshgrid(np.arange(sz) - (sz - 1) / 2, np.arange(sy) - (sy - 1) / 2,
                                 np.arange(sx) - (sx - 1) / 2, indexing="ij")
        r = np.stack([lx / max(sx / 2, 1), ly / max(sy / 2, 1), lz / max(sz / 2, 1)], axis=-1)
        axis = np.argmax(np.abs(r), axis=-1)[..., None]
        comp = np.take_along_axis(r, axis, -1)
        n = np.zeros(r.shape, np.float32)
        np.put_along_axis(n, axis, np.where(comp < 0, -1.0, 1.0).astype(np.float32), -1)
        normals[sl] = n

    if n_materials >= 2 and emissive_fraction > 0:
        zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij") if X * Y * Z <= 1 << 22 \
            else (None, None, None)
        if zz is not None:
            em = hash_u32(xx, yy, zz, seed=seed + 4) < np.uint32(emissive_fraction * 4294967296.0)
            mat[em & solid] = n_materials - 1
        else:
            for z in range(Z):
                yy2, xx2 = np.meshgrid(np.arange(Y), np.arange(X), indexing="ij")
                em = hash_u32(xx2, yy2, np.full_like(xx2, z), seed=seed + 4) < np.uint32(emissive_fraction * 4294967296.0)
                mat[z][em & solid[z]] = n_materials - 1

    mats = _materials(n_materials, seed + 1, n_materials - 1 if n_materials >= 2 else None, emission)
    mat = np.where(solid, mat, EMPTY).astype(np.uint32)
    normal_words = pack_normal_word(mat, normals)
    normal_words[~solid] = 0xFFFFFFFF
    alb = np.array([m.albedo for m in mats] + [(0, 0, 0)] * (256 - len(mats)), dtype=np.uint32)
    mi = np.minimum(mat, 255)
    albedo_words = (alb[mi, 0] << 24) | (alb[mi, 1] << 16) | (alb[mi, 2] << 8)
    albedo_words[~solid] = 0
    return normal_words, albedo_words.astype(np.uint32), mats


def config_small(seed: int = 11, chunks=(4, 3, 4), occupancy: float = 0.3, n_boxes: int = 3, reference_like=True,
                 width: int = 48, height: int = 40) -> Scene:
    """A few chunks: the size the CPU oracle checks in seconds."""
    X, Y, Z = (c * 8 for c in chunks)
    nw, aw, mats = terrain((X, Y, Z), occupancy, seed=seed, amplitude=0.3 * Y, n_boxes=n_boxes, box_size=(2, 9),
                           n_materials=4, emissive_fraction=0.02, emission=0.0 if reference_like else 4.0)
    return Scene("small", tuple(chunks), nw, aw, mats, cam_pos=(0.5, chunks[1] * 0.87, 0.5),
                 cam_orient=(50.0, 45.0, 0.0), width=width, height=height)


def config1(reference_like: bool = False, emission: float | None = None) -> Scene:
    """16x16x16 chunks (128^3 voxels), fBm terrain at ~35% occupancy, 4 materials
    (one emissive, strength 4 unless `emission` is given), 256x256 camera;
    16 direct/specular rays, 1 diffuse sample."""
    em = emission if emission is not None else (0.0 if reference_like else 4.0)
    nw, aw, mats = terrain((128, 128, 128), 0.35, seed=1, amplitude=48.0, n_materials=4, emission=em)
    params = {} if reference_like else {"numSpecularRays": 16, "numDirectRays": 16}
    return Scene("config1", (16, 16, 16), nw, aw, mats, cam_pos=(1.0, 8.0, 1.0), cam_orient=(35.0, 45.0, 0.0),
                 width=256, height=256, params=params)


def config2() -> Scene:
    """64x32x64 chunks (512x256x512 voxels), fBm terrain + 2000 boxes (sizes 2-40),
    ~30% occupancy; 1920x1080 draw; 32 direct/specular rays, <= 4 diffuse bounces."""
    nw, aw, mats = terrain((512, 256, 512), 0.22, seed=1, amplitude=64.0, n_boxes=2000, box_seed=3,
                           box_size=(2, 40), n_materials=4, emission=4.0)
    params = {"numSpecularRays": 32, "numDirectRays": 32, "diffuseBounceLimit": 4}
    return Scene("config2", (64, 32, 64), nw, aw, mats, cam_pos=(4.0, 12.0, 4.0), cam_orient=(30.0, 45.0, 0.0),
                 width=1920, height=1080, params=params)


def batch_scene(index: int, max_grid: int = 48) -> Scene:
    """Scene `index` of the batched uneven workload (seed = index): chunk grid
    uniform in [4, max_grid] per axis, fBm heightfield at an occupancy uniform in
    [5%, 60%], 2-16 materials (the last one emissive, on ~0.5% of columns), and
    a visible-chunk budget log-uniform in [10, 40000] (capped by the grid later
    by the exposed-chunk count). Voxels of a column share the column's normal
    and material, so the volume is built from 2D arrays."""
    def u(k):
        return float(uniform01(index, k, seed=977))
    span = max_grid - 4 + 1
    gx, gy, gz = (4 + min(span - 1, int(span * u(k))) for k in range(3))
    occ = 0.05 + 0.55 * u(3)
    n_mat = 2 + min(14, int(15 * u(4)))
    visible = int(round(np.exp(np.log(10.0) + (np.log(40000.0) - np.log(10.0)) * u(5))))
    X, Y, Z = gx * 8, gy * 8, gz * 8
    gxs, gzs = np.meshgrid(np.arange(X, dtype=np.float64), np.arange(Z, dtype=np.float64))
    f = fbm(gxs / 48.0, gzs / 48.0, seed=index * 7 + 1)
    amp = 0.3 * Y
    lo, hi = -amp, float(Y)
    for _ in range(50):
        h0 = 0.5 * (lo + hi)
        if np.clip(h0 + amp * f, 0, Y).mean() / Y < occ:
            lo = h0
        else:
            hi = h0
    h = h0 + amp * f
    hz, hx = np.gradient(h)
    nrm = np.stack([-hx, np.ones_like(h), -hz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    ix, iz = np.meshgrid(np.arange(X), np.arange(Z))
    mat = hash_u32(ix, iz, seed=index + 5) % np.uint32(max(1, n_mat - 1))
    emissive = hash_u32(ix, iz, seed=index + 6) < np.uint32(0.005 * 4294967296.0)
    mat = np.where(emissive, n_mat - 1, mat).astype(np.uint32)
    mats = _materials(n_mat, index + 11, n_mat - 1, 1.0)
    col_nw = pack_normal_word(mat, nrm)                              # (Z, X)
    alb = np.array([m.albedo for m in mats], dtype=np.uint32)
    col_aw = (alb[mat, 0] << 24) | (alb[mat, 1] << 16) | (alb[mat, 2] << 8)
    solid = np.arange(Y)[None, :, None] < h[:, None, :]              # (Z, Y, X)
    nw = np.where(solid, col_nw[:, None, :], np.uint32(0xFFFFFFFF)).astype(np.uint32)
    aw = np.where(solid, col_aw[:, None, :], np.uint32(0)).astype(np.uint32)
    params = {"numSpecularRays": 16, "numDirectRays": 16, "diffuseBounceLimit": 4}
    return Scene(f"batch{index}", (gx, gy, gz), nw, aw, mats, cam_pos=(gx * 0.5, gy * 0.9, gz * 0.5),
                 cam_orient=(45.0, 45.0, 0.0), params=params, visible=visible)


def sphere_edits(scene: Scene, frame: int, n: int = 256) -> list:
    """The dynamic-edit workload's edits of one frame: `n` spheres of radius
    U[3, 24] voxels, centre uniform over the map, carve or fill 50/50 (a fill
    takes one of the scene's non-emissive materials and its albedo), seed
    4 + frame. Each edit is (material, albedo, centre, radius, flip_normals)."""
    X, Y, Z = (8 * c for c in scene.map_size)
    n_mats = len(scene.materials) - 1
    kk, jj = np.meshgrid(np.arange(n), np.arange(6), indexing="ij")
    us = uniform01(kk, jj, seed=4 + frame).tolist()
    out = []
    for u in us:
        r = 3.0 + 21.0 * u[0]
        ctr = (u[1] * X, u[2] * Y, u[3] * Z)
        if u[4] < 0.5:
            out.append((255, (0, 0, 0), ctr, r, False))
        else:
            m = min(n_mats - 1, int(u[5] * n_mats))
            out.append((m, scene.materials[m].albedo, ctr, r, False))
    return out

"""The oracle's restatement of the reference's `.voxvol` volume file format
(test infrastructure: only tests/ use it, as the checker).

Follows src/DoonEngine/voxel.c:
  _DN_compress_chunk   :300-430   per-chunk encoder (palettes, material runs)
  _DN_decompress_chunk :433-518   per-chunk decoder
  DN_load_volume       :520-593   file layout, read side
  DN_save_volume       :595-654   file layout, write side

File layout (little endian, x86-64 sizes): map size (3 x u32), chunk capacity
(size_t, u64), then per chunk slot a u16 byte count and the compressed chunk,
then 256 materials of 32 bytes, camera (pos 3f, orient 3f, fov f, view mode
u32), lighting (sun dir 3f, sun strength 3f, ambient 3f, diffuse bounce limit
u32, specular bounce limit u32, shadow softness f) and sky (bottom 3f, top 3f).

Compressed chunk: its position (3 x i32); nothing else when that position is
outside the map (an unused slot). Otherwise a normal palette and an albedo
palette (u8 count, then count x 3 bytes; count 0 = no palette), then runs over
the 512 voxels in x-fastest order: material (u8), run length (u8, at most 255),
and for each solid voxel of the run its normal (palette index u8, or 3 bytes
x, y, z of the packed normal) and its albedo (palette index, or 3 bytes r, g,
b). A palette is written only while it stays smaller than half the chunk's
stored-voxel count (integer division), entries in first-seen order.

Voxels are (512, 2) u32 arrays of (normal word, albedo word), index
x + 8*y + 64*z here (the file's order); empty voxels decode as normal
0xFFFFFFFF with the albedo left as it was (zero in a fresh chunk).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

EMPTY = 255
TAIL = struct.Struct("<3f3ffI3f3f3fIIf3f3f")   # camera, lighting, sky (voxel.c:577-591)


@dataclass
class Chunk:
    pos: tuple
    voxels: np.ndarray | None = None   # (512, 2) u32, None for an unused slot
    num_voxels: int = 0


@dataclass
class VolumeFile:
    map_size: tuple
    chunks: list = field(default_factory=list)
    materials: bytes = b""
    tail: tuple = ()


def in_map(pos, map_size) -> bool:
    return all(0 <= p < m for p, m in zip(pos, map_size))


def decode_chunk(mem: bytes, map_size) -> tuple[Chunk, int]:
    """_DN_decompress_chunk (voxel.c:433-518). Returns the chunk and the bytes used."""
    pos = struct.unpack_from("<3i", mem, 0)
    q = 12
    if not in_map(pos, map_size):
        return Chunk(pos), q
    nn = mem[q]; q += 1
    npal = [tuple(mem[q + 3 * i: q + 3 * i + 3]) for i in range(nn)]; q += 3 * nn
    na = mem[q]; q += 1
    apal = [tuple(mem[q + 3 * i: q + 3 * i + 3]) for i in range(na)]; q += 3 * na
    vox = np.zeros((512, 2), np.uint32)
    count = 0
    read = 0
    while read < 512:
        mat, num = mem[q], mem[q + 1]
        q += 2
        for i in range(read, read + num):
            if mat == EMPTY:
                vox[i, 0] = 0xFFFFFFFF
                continue
            if nn:
                n = npal[mem[q]]; q += 1
            else:
                n = tuple(mem[q:q + 3]); q += 3
            if na:
                a = apal[mem[q]]; q += 1
            else:
                a = tuple(mem[q:q + 3]); q += 3
            vox[i, 0] = (mat << 24) | (n[0] << 16) | (n[1] << 8) | n[2]
            vox[i, 1] = (a[0] << 24) | (a[1] << 16) | (a[2] << 8)
            count += 1
        read += num
    return Chunk(pos, vox, count), q


def encode_chunk(ch: Chunk, map_size) -> bytes:
    """_DN_compress_chunk (voxel.c:300-430)."""
    out = bytearray(struct.pack("<3i", *ch.pos))
    if not in_map(ch.pos, map_size):
        return bytes(out)
    vox = ch.voxels
    half = ch.num_voxels // 2
    npal, apal = [], []
    for i in range(512):
        w = int(vox[i, 0])
        if (w >> 24) == EMPTY:
            continue
        n = ((w >> 16) & 255, (w >> 8) & 255, w & 255)
        a = int(vox[i, 1])
        a = ((a >> 24) & 255, (a >> 16) & 255, (a >> 8) & 255)
        if len(npal) < half and n not in npal:
            npal.append(n)
        if len(apal) < half and a not in apal:
            apal.append(a)
    if not len(npal) < half:
        npal = []
    if not len(apal) < half:
        apal = []
    out.append(len(npal))
    for n in npal:
        out += bytes(n)
    out.append(len(apal))
    for a in apal:
        out += bytes(a)
    i = 0
    while i < 512:
        mat = int(vox[i, 0]) >> 24
        out.append(mat)
        at = len(out)
        out.append(0)
        num = 0
        j = i
        while j < 512 and num < 255 and (int(vox[j, 0]) >> 24) == mat:
            num += 1
            if mat != EMPTY:
                w = int(vox[j, 0])
                n = ((w >> 16) & 255, (w >> 8) & 255, w & 255)
                out += bytes([npal.index(n)]) if npal else bytes(n)
                a = int(vox[j, 1])
                a = ((a >> 24) & 255, (a >> 16) & 255, (a >> 8) & 255)
                out += bytes([apal.index(a)]) if apal else bytes(a)
            j += 1
        out[at] = num
        i = j
    return bytes(out)


def decode(data: bytes) -> VolumeFile:
    """DN_load_volume's reading (voxel.c:520-593); raises on a truncated file."""
    map_size = struct.unpack_from("<3I", data, 0)
    cap, = struct.unpack_from("<Q", data, 12)
    q = 20
    vf = VolumeFile(map_size)
    for _ in range(cap):
        size, = struct.unpack_from("<H", data, q)
        q += 2
        ch, used = decode_chunk(data[q:q + size], map_size)
        if used != size:
            raise ValueError(f"chunk record of {size} bytes decodes {used}")
        vf.chunks.append(ch)
        q += size
    vf.materials = bytes(data[q:q + 256 * 32])
    q += 256 * 32
    vf.tail = TAIL.unpack_from(data, q)
    q += TAIL.size
    if q != len(data):
        raise ValueError(f"{len(data) - q} trailing bytes")
    return vf


def encode(vf: VolumeFile) -> bytes:
    """DN_save_volume's writing (voxel.c:595-654)."""
    out = bytearray(struct.pack("<3IQ", *vf.map_size, len(vf.chunks)))
    for ch in vf.chunks:
        rec = encode_chunk(ch, vf.map_size)
        out += struct.pack("<H", len(rec)) + rec
    out += vf.materials
    out += TAIL.pack(*vf.tail)
    return bytes(out)


def chunk_array(vox: np.ndarray) -> np.ndarray:
    """(512, 2) in file order (x fastest) -> (512, 2) in the CPU chunk's
    voxels[x][y][z] order (index 64x + 8y + z)."""
    return vox.reshape(8, 8, 8, 2).transpose(2, 1, 0, 3).reshape(512, 2)

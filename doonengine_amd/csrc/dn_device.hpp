// Device-side voxel traversal for gfx950: the two-level DDA of the reference
// (assets/shaders/voxelShared.comp:300-475) with its codec and rank/select.
//
// Floating-point semantics are fixed so results are reproducible bit for bit
// (the translation unit is built with -ffp-contract=off, correctly rounded
// division and sqrt, IEEE denormals):
//   min(x,y) = y<x ? y : x   max(x,y) = x<y ? y : x   sign(0) = 0
//   dot = (x*x' + y*y') + z*z'   normalize(v) = v * (1/sqrt(dot(v,v)))
//   round = round-half-to-even
#pragma once
#include "dn_internal.h"

#define DN_EPS 0.0001f
#define DN_INV255 0.00392156862f      // voxelShared.comp:249 (one ulp below 1/255)
#define DN_INV65535 0.0000152590219f  // voxelShared.comp:252

struct f3 { float x, y, z; };

__device__ __forceinline__ f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3 splat(float s) { return mk(s, s, s); }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator+(f3 a, float s) { return mk(a.x + s, a.y + s, a.z + s); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ f3 vfloor(f3 a) { return mk(floorf(a.x), floorf(a.y), floorf(a.z)); }
__device__ __forceinline__ f3 vabs(f3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__device__ __forceinline__ f3 vinv(f3 a) { return mk(1.0f / a.x, 1.0f / a.y, 1.0f / a.z); }
__device__ __forceinline__ f3 vmin(f3 a, f3 b) { return mk(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
__device__ __forceinline__ f3 vmax(f3 a, f3 b) { return mk(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
__device__ __forceinline__ f3 vclamp(f3 a, float lo, float hi)
{
    return mk(gmin(gmax(a.x, lo), hi), gmin(gmax(a.y, lo), hi), gmin(gmax(a.z, lo), hi));
}
__device__ __forceinline__ f3 vnorm(f3 a) { float inv = 1.0f / sqrtf(dot3(a, a)); return a * inv; }
__device__ __forceinline__ f3 vreflect(f3 I, f3 N) { return I - N * (2.0f * dot3(N, I)); }
__device__ __forceinline__ f3 vrefract(f3 I, f3 N, float eta)
{
    float d = dot3(N, I);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if(k < 0.0f) return splat(0.0f);
    return I * eta - N * (eta * d + sqrtf(k));
}
__device__ __forceinline__ f3 vmix(f3 x, f3 y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float minc(f3 a) { return gmin(gmin(a.x, a.y), a.z); }
__device__ __forceinline__ f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ f3 vpow(f3 a, float e) { return mk(powf(a.x, e), powf(a.y, e), powf(a.z, e)); }

struct DVoxel { f3 normal; uint32_t material; f3 albedo; f3 spec; f3 diffuse; };

// decompress_voxel, voxelShared.comp:236-256
__device__ __forceinline__ DVoxel decompress(uint4 w)
{
    DVoxel r;
    uint32_t n1 = (w.x >> 16) & 255u, n2 = (w.x >> 8) & 255u, n3 = w.x & 255u;
    uint32_t a0 = w.y >> 24, a1 = (w.y >> 16) & 255u, a2 = (w.y >> 8) & 255u, a3 = w.y & 255u;
    uint32_t s0 = w.z >> 24, s1 = (w.z >> 16) & 255u;
    uint32_t dx = w.z & 0xFFFFu, dy = w.w >> 16, dz = w.w & 0xFFFFu;
    r.normal = (mk((float)n1, (float)n2, (float)n3) * DN_INV255 - splat(0.5f)) * 2.0f;
    r.material = w.x >> 24;
    r.albedo = mk((float)a0, (float)a1, (float)a2) * DN_INV255;
    r.diffuse = mk((float)dx, (float)dy, (float)dz) * DN_INV65535;
    r.spec = mk((float)a3, (float)s0, (float)s1) * DN_INV255;
    return r;
}

__device__ __forceinline__ float emission_of(const DNmat& m) { return m.padding[0] > 0.0f ? m.padding[0] : 1.0f; }

// ---------------------------------------------------------------------------
// occupancy mask held in registers: 16 dwords as 4 x uint4

struct Mask512 { uint4 q[4]; };

__device__ __forceinline__ uint4 sel_quarter(const Mask512& m, uint32_t q)
{
    uint4 lo = (q & 1u) ? m.q[1] : m.q[0];
    uint4 hi = (q & 1u) ? m.q[3] : m.q[2];
    return (q & 2u) ? hi : lo;
}
__device__ __forceinline__ uint32_t sel_word(uint4 v, uint32_t j)
{
    uint32_t lo = (j & 1u) ? v.y : v.x;
    uint32_t hi = (j & 1u) ? v.w : v.z;
    return (j & 2u) ? hi : lo;
}
__device__ __forceinline__ uint32_t mask_bit(const Mask512& m, uint32_t li)
{
    uint32_t w = li >> 5;
    return (sel_word(sel_quarter(m, w >> 2), w & 3u) >> (li & 31u)) & 1u;
}

// get_voxel_index without the pool base (voxelShared.comp:150-169)
__device__ __forceinline__ uint32_t rank_in(const Mask512& m, const uint32_t partial[3], uint32_t li)
{
    uint32_t w = li >> 5, q = w >> 2, j = w & 3u;
    uint32_t n = q == 0 ? 0u : (q == 1 ? partial[0] : (q == 2 ? partial[1] : partial[2]));
    uint4 v = sel_quarter(m, q);
    uint32_t part = (1u << (li & 31u)) - 1u;
    uint32_t b0 = j == 0 ? (v.x & part) : v.x;
    uint32_t b1 = j == 1 ? (v.y & part) : (j > 1 ? v.y : 0u);
    uint32_t b2 = j == 2 ? (v.z & part) : (j > 2 ? v.z : 0u);
    uint32_t b3 = j == 3 ? (v.w & part) : 0u;
    return n + __builtin_popcount(b0) + __builtin_popcount(b1) + __builtin_popcount(b2) + __builtin_popcount(b3);
}

// get_voxel_position (voxelShared.comp:172-224): position index of the
// (voxNum+1)-th stored voxel, or -1.
// Position of the k-th (from 0) set bit of m, which has more than k set bits:
// a binary search over popcounts of halves, quarters, ... of the word.
__device__ __forceinline__ uint32_t select_bit32(uint32_t m, uint32_t k)
{
    uint32_t pos = 0, c;
    c = __builtin_popcount(m & 0xFFFFu); if(k >= c) { k -= c; m >>= 16; pos += 16; }
    c = __builtin_popcount(m & 0xFFu);   if(k >= c) { k -= c; m >>= 8; pos += 8; }
    c = __builtin_popcount(m & 0xFu);    if(k >= c) { k -= c; m >>= 4; pos += 4; }
    c = __builtin_popcount(m & 0x3u);    if(k >= c) { k -= c; m >>= 2; pos += 2; }
    if(k >= (m & 1u)) pos += 1;
    return pos;
}

// get_voxel_position's select (voxelShared.comp:172-224): the bit index of the
// voxNum-th (from 0) stored voxel of a chunk, or -1 past the last one. `mask`
// is a chunk record's 16-word mask (16-byte aligned): its four 16-byte pieces
// are loaded at once, then the word holding the bit is found by popcounts.
__device__ __forceinline__ int32_t select_in(const uint32_t* mask, uint32_t voxNum)
{
    const uint4* q = reinterpret_cast<const uint4*>(mask);
    const uint4 a = q[0], b = q[1], c4 = q[2], d = q[3];
    const uint32_t w[16] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c4.x, c4.y, c4.z, c4.w, d.x, d.y, d.z, d.w};
    uint32_t count = 0;
    int32_t result = -1;
    #pragma unroll
    for(uint32_t i = 0; i < 16; i++)
    {
        const uint32_t bc = (uint32_t)__builtin_popcount(w[i]);
        if(result < 0 && count + bc > voxNum) result = (int32_t)(i * 32u + select_bit32(w[i], voxNum - count));
        count += bc;
    }
    return result;
}

// ---------------------------------------------------------------------------
// two-level DDA

struct DDA { int32_t px, py, pz; f3 delta; int32_t sx, sy, sz; f3 side; };

__device__ __forceinline__ void init_dda(f3 dir, f3 inv, f3 p, DDA& d)
{
    f3 f = vfloor(p);
    d.px = (int32_t)f.x; d.py = (int32_t)f.y; d.pz = (int32_t)f.z;
    d.delta = vabs(inv);
    f3 s = mk(gsign(dir.x), gsign(dir.y), gsign(dir.z));
    d.sx = (int32_t)s.x; d.sy = (int32_t)s.y; d.sz = (int32_t)s.z;
    f3 fp = mk((float)d.px, (float)d.py, (float)d.pz);
    d.side = ((s * (fp - p) + s * 0.5f) + 0.5f) * d.delta;
}

__device__ __forceinline__ void iterate_dda(DDA& d, f3& normal)
{
    f3 sd = d.side;
    bool mx = sd.x <= gmin(sd.y, sd.z);
    bool my = sd.y <= gmin(sd.z, sd.x);
    bool mz = sd.z <= gmin(sd.x, sd.y);
    float fx = mx ? 1.0f : 0.0f, fy = my ? 1.0f : 0.0f, fz = mz ? 1.0f : 0.0f;
    d.side.x = sd.x + fx * d.delta.x;
    d.side.y = sd.y + fy * d.delta.y;
    d.side.z = sd.z + fz * d.delta.z;
    d.px += mx ? d.sx : 0; d.py += my ? d.sy : 0; d.pz += mz ? d.sz : 0;
    normal.x = fx * (float)(-d.sx);
    normal.y = fy * (float)(-d.sy);
    normal.z = fz * (float)(-d.sz);
}

// iterate_dda for a ray whose delta = |1/dir| is finite on every axis (no
// direction component is exactly zero). Every sideDist is then finite and
// >= +0, so GLSL's min (y < x ? y : x) is the hardware min, and
// side + mask * delta with mask in {0, 1} is side + delta or side itself
// (0 * finite delta = +0 and side + +0 = side): the same values as iterate_dda
// in fewer operations. With a zero direction component delta is infinite and
// the reference's 0 * delta is NaN, which its arithmetic carries on; such rays
// keep iterate_dda.
// The side distances are also never -0 (the start value is a product of two
// non-negative terms whose first factor cancels to +0, not -0), so their bit
// patterns order as unsigned integers exactly as the values do: the minima and
// comparisons below are integer operations on the bits.
__device__ __forceinline__ void iterate_dda_finite(DDA& d, f3& normal)
{
    f3 sd = d.side;
    const uint32_t bx = __float_as_uint(sd.x), by = __float_as_uint(sd.y), bz = __float_as_uint(sd.z);
    bool mx = bx <= min(by, bz);
    bool my = by <= min(bz, bx);
    bool mz = bz <= min(bx, by);
    d.side.x = mx ? sd.x + d.delta.x : sd.x;
    d.side.y = my ? sd.y + d.delta.y : sd.y;
    d.side.z = mz ? sd.z + d.delta.z : sd.z;
    d.px += mx ? d.sx : 0; d.py += my ? d.sy : 0; d.pz += mz ? d.sz : 0;
    normal.x = (mx ? 1.0f : 0.0f) * (float)(-d.sx);
    normal.y = (my ? 1.0f : 0.0f) * (float)(-d.sy);
    normal.z = (mz ? 1.0f : 0.0f) * (float)(-d.sz);
}

__device__ __forceinline__ bool all_finite(f3 v)
{
    return fabsf(v.x) < INFINITY && fabsf(v.y) < INFINITY && fabsf(v.z) < INFINITY;
}

// Cell tests as unsigned compares: a negative coordinate wraps above any bound.
__device__ __forceinline__ bool in_chunk_cell(const DDA& d)
{
    return max(max((uint32_t)d.px, (uint32_t)d.py), (uint32_t)d.pz) < 8u;
}

__device__ __forceinline__ bool in_map_cell(const DDA& d, uint32_t X, uint32_t Y, uint32_t Z)
{
    return (uint32_t)d.px < X && (uint32_t)d.py < Y && (uint32_t)d.pz < Z;
}

// Per-ray mutable state that the shaders keep in invocation globals
// (voxelShared.comp:321-325), plus per-lane counters.
struct RayCtx {
    uint32_t lastVoxID;
    float lastVoxRefract;
    uint32_t hitIndex;
    uint32_t sMap, sVox;          // trips of the last step_map
    uint32_t cRays, cMap, cVox, cHits;
    uint32_t* maskRow;            // this lane's LDS row for the chunk mask (LDS-mask walks)
    const uint32_t* sharedRow;    // an LDS row already holding the mask of tile sharedMap
    uint32_t sharedMap;           // (0xFFFFFFFF: none)
};

// LDS rows for the chunk masks of the lanes of a block: 17 words per lane (16
// mask words + 1 of padding), so the same word index of 64 lanes falls in 64
// different banks and a voxel trip reads its mask word conflict-free.
#define DN_MASK_ROW 17u
#define DN_MASK_LDS_WORDS(threads) ((threads) * DN_MASK_ROW)

__device__ __forceinline__ void ctx_init(RayCtx& c)
{
    c.maskRow = nullptr;
    c.sharedRow = nullptr;
    c.sharedMap = 0xFFFFFFFFu;
    c.lastVoxID = 255u; c.lastVoxRefract = 1.0f; c.hitIndex = 0xFFFFFFFFu;
    c.sMap = c.sVox = 0; c.cRays = c.cMap = c.cVox = c.cHits = 0;
}

__device__ __forceinline__ bool occ_bit(const uint32_t* bits, uint32_t idx)
{
    return (bits[idx >> 5] >> (idx & 31u)) & 1u;
}

// rank_in over a mask held in an LDS row (same arithmetic)
__device__ __forceinline__ uint32_t rank_row(const uint32_t* row, const uint32_t partial[3], uint32_t li)
{
    uint32_t w = li >> 5, q = w >> 2, j = w & 3u;
    uint32_t n = q == 0 ? 0u : (q == 1 ? partial[0] : (q == 2 ? partial[1] : partial[2]));
    const uint32_t* v = row + 4u * q;
    uint32_t part = (1u << (li & 31u)) - 1u;
    uint32_t b0 = j == 0 ? (v[0] & part) : v[0];
    uint32_t b1 = j == 1 ? (v[1] & part) : (j > 1 ? v[1] : 0u);
    uint32_t b2 = j == 2 ? (v[2] & part) : (j > 2 ? v[2] : 0u);
    uint32_t b3 = j == 3 ? (v[3] & part) : 0u;
    return n + __builtin_popcount(b0) + __builtin_popcount(b1) + __builtin_popcount(b2) + __builtin_popcount(b3);
}

// step_chunk, voxelShared.comp:328-418. maxDepth is negative on every path here
// (no raster compositing), so the depth test at :363 always passes.
// LDSM: the chunk's occupancy mask goes to the lane's LDS row (c.maskRow) on
// entry and each voxel trip reads one word of it, instead of selecting the
// word out of 16 registers by a per-lane index. FINITE: the ray's delta is
// finite on every axis, so iterate_dda_finite applies (never with refraction,
// which changes the direction).
// lastSideDist of voxelShared.comp: the side distances before the last step,
// of which the walk only ever uses the minimum (the distance at which the ray
// entered its current cell). The finite-delta walk keeps just that minimum
// (min is exact and order-free on finite values, and on their bits, see
// iterate_dda_finite); the reference walk keeps the three distances and takes
// GLSL's min when asked.
template <bool FINITE> struct LastSide {
    f3 v = splat(0.0f);
    __device__ __forceinline__ void set(f3 s) { v = s; }
    __device__ __forceinline__ void reset() { v = splat(0.0f); }
    __device__ __forceinline__ float minc() const { return ::minc(v); }
};
template <> struct LastSide<true> {
    float t = 0.0f;
    __device__ __forceinline__ void set(f3 s)
    {
        t = __uint_as_float(min(min(__float_as_uint(s.x), __float_as_uint(s.y)), __float_as_uint(s.z)));
    }
    __device__ __forceinline__ void reset() { t = 0.0f; }
    __device__ __forceinline__ float minc() const { return t; }
};

template <bool REFRACT, bool LDSM = false, bool FINITE = false>
__device__ bool step_chunk(const DNsceneView& s, RayCtx& c, uint32_t mapIndex, f3& dir, f3& inv, f3& p,
                           bool ignoreFirst, f3& hitNormal, DVoxel& vox, f3& colorAdd, float& colorMult,
                           bool& refracted)
{
    refracted = false;
    const DNchunkRec* rec = s.chunks + mapIndex;
    const uint4* rq = reinterpret_cast<const uint4*>(rec);
    Mask512 m;
    const uint32_t* mrow = c.maskRow;
    if(LDSM)
    {
        if(mapIndex == c.sharedMap)
            mrow = c.sharedRow;            // staged once for the whole wave
        else
        {
            uint32_t* row = c.maskRow;
            for(int k = 0; k < 4; k++)
            {
                uint4 v = rq[k];
                row[4 * k + 0] = v.x; row[4 * k + 1] = v.y; row[4 * k + 2] = v.z; row[4 * k + 3] = v.w;
            }
        }
    }
    else
    {
        m.q[0] = rq[0]; m.q[1] = rq[1]; m.q[2] = rq[2]; m.q[3] = rq[3];
    }
    DDA d;
    init_dda(dir, inv, p, d);
    LastSide<FINITE> lsd;
    static_assert(!(REFRACT && FINITE), "refraction changes the direction: no finite-delta walk");
    while(in_chunk_cell(d))
    {
        c.sVox++;
        uint32_t li = (uint32_t)(d.px + 8 * (d.py + 8 * d.pz));
        const uint32_t bit = LDSM ? (mrow[li >> 5] >> (li & 31u)) & 1u : mask_bit(m, li);
        if(bit && !ignoreFirst)
        {
            uint4 tail = rq[4];   // partial[0..2], voxelIndex
            uint32_t partial[3] = {tail.x, tail.y, tail.z};
            uint32_t idx = tail.w + (LDSM ? rank_row(mrow, partial, li) : rank_in(m, partial, li));
            uint4 cw = s.voxels[idx];
            vox = decompress(cw);
            const DNmat& mat = s.mats[vox.material];
            uint32_t thisID = (cw.y & 0xFFFFFF00u) | (cw.x >> 24);
            if(mat.opacity == 1.0f)
            {
                p = p + dir * (lsd.minc() + DN_EPS);
                c.hitIndex = idx;
                return true;
            }
            else if(c.lastVoxID != thisID)
            {
                float op = mat.opacity;
                colorAdd = colorAdd + (vox.albedo * (colorMult * op)) * ld3(s.sunStrength);
                colorMult = colorMult * (1.0f - op);
                float rayDist = lsd.minc();
                if(REFRACT && rayDist > 0.0f)
                {
                    refracted = true;
                    p = p + dir * (rayDist + DN_EPS);
                    f3 n = dot3(vox.normal, dir) < 0.0f ? vnorm(vox.normal) : hitNormal;
                    dir = vrefract(dir, n, c.lastVoxRefract / mat.refractIndex);
                    inv = vinv(dir);
                    init_dda(dir, inv, p, d);
                    lsd.reset();
                }
                c.lastVoxID = thisID;
                c.lastVoxRefract = mat.refractIndex;
            }
        }
        else if(c.lastVoxID != 255u)
        {
            if(REFRACT)
            {
                refracted = true;
                p = p + dir * (lsd.minc() + DN_EPS);
                f3 old = dir;
                f3 nn = -vox.normal;
                f3 n = dot3(nn, dir) < 0.0f ? vnorm(nn) : hitNormal;
                dir = vrefract(dir, n, c.lastVoxRefract);
                if(dir.x == 0.0f && dir.y == 0.0f && dir.z == 0.0f) dir = old;
                inv = vinv(dir);
                init_dda(dir, inv, p, d);
                lsd.reset();
            }
            c.lastVoxID = 255u;
            c.lastVoxRefract = 1.0f;
        }
        lsd.set(d.side);
        if(FINITE) iterate_dda_finite(d, hitNormal);
        else iterate_dda(d, hitNormal);
        ignoreFirst = false;
    }
    return false;
}

// step_map, voxelShared.comp:421-475
template <bool REFRACT, bool LDSM = false, bool FINITE = false>
__device__ bool step_map(const DNsceneView& s, RayCtx& c, f3& dir, f3 inv, f3& p, bool ignoreFirst, f3& hitNormal,
                         DVoxel& vox, f3& colorAdd, float& colorMult)
{
    bool requested = false;
    colorAdd = splat(0.0f);
    colorMult = 1.0f;
    c.sMap = 0; c.sVox = 0; c.hitIndex = 0xFFFFFFFFu;
    DDA d;
    init_dda(dir, inv, p, d);
    LastSide<FINITE> lsd;
    bool hit = false;
    const int32_t X = (int32_t)s.mapX, Y = (int32_t)s.mapY, Z = (int32_t)s.mapZ;
    // a lighting ray that does not move down reads the column table of its
    // quadrant: the highest present tile over the columns it can still reach
    const bool colRule = !REFRACT && s.colTop != nullptr && d.sy >= 0;
    const uint32_t colBase = ((d.sx > 0 ? 1u : 0u) + (d.sz > 0 ? 2u : 0u)) * s.mapX * s.mapZ;
    while(in_map_cell(d, (uint32_t)X, (uint32_t)Y, (uint32_t)Z))
    {
        c.sMap++;
        const int32_t colTop = colRule ? s.colTop[colBase + (uint32_t)d.px + s.mapX * (uint32_t)d.pz] : INT32_MAX;
        uint32_t idx = (uint32_t)d.px + s.mapX * ((uint32_t)d.py + s.mapY * (uint32_t)d.pz);
        if(occ_bit(s.occLoaded, idx))
        {
            s.lastUsed[idx] = 0u;          // a plain store: no read on the ray's path
            f3 upd = p + dir * (lsd.minc() - DN_EPS);
            f3 crp = (upd - mk((float)d.px, (float)d.py, (float)d.pz)) * 8.0f;
            crp = vclamp(crp, DN_EPS, 8.0f - DN_EPS);
            bool refr;
            if(step_chunk<REFRACT, LDSM, FINITE>(s, c, idx, dir, inv, crp, ignoreFirst, hitNormal, vox, colorAdd, colorMult,
                                         refr))
            {
                p = mk((float)d.px, (float)d.py, (float)d.pz) + crp * 0.125f;
                hit = true;
                break;
            }
            if(REFRACT && refr)
            {
                p = mk((float)d.px, (float)d.py, (float)d.pz) + crp * 0.125f;
                init_dda(dir, inv, p, d);
                lsd.reset();
            }
        }
        else if(s.hasUnloaded && !requested && occ_bit(s.occPresent, idx))
        {
            s.flags[idx] = 3u;
            requested = true;
        }
        lsd.set(d.side);
        if(FINITE) iterate_dda_finite(d, hitNormal);
        else iterate_dda(d, hitNormal);
        ignoreFirst = false;
        // A lighting ray beyond the box of present tiles on the side it moves
        // toward (a positive step past presHi, any other step below presLo)
        // meets no loaded or unloaded tile before it leaves the map, so the
        // rest of the walk would only count trips: it ends here as the same
        // miss, with the same colour terms. So does a ray that does not move
        // down once it is above every present tile of the columns it could
        // still reach from the cell it just left (its y only grows and those
        // columns include all of its later ones). Only the trip counters
        // differ from the reference walk's; the oracle applies the same rules.
        if(!REFRACT && ((d.sx > 0 ? d.px > s.presHi[0] : d.px < s.presLo[0]) ||
                        (d.sy > 0 ? d.py > s.presHi[1] : d.py < s.presLo[1]) ||
                        (d.sz > 0 ? d.pz > s.presHi[2] : d.pz < s.presLo[2]) || d.py > colTop))
            break;
    }
    c.cRays++;
    c.cMap += c.sMap;
    c.cVox += c.sVox;
    c.cHits += hit ? 1u : 0u;
    return hit;
}

// step_map without refraction, as one loop over both DDA levels. The chunk
// level and the map level share delta = |1/dir| and step = sign(dir), so the
// trip ends with the same `lsd = side; iterate` for every lane whatever its
// level, and only the occupancy test differs: lanes of a wave stay converged
// instead of waiting on each other's inner chunk loops. The sequence of
// floating-point operations per ray is exactly that of step_map<false>
// (voxelShared.comp:328-475): map state is saved on chunk entry and resumed,
// then stepped, when the chunk walk leaves the chunk.
// With NORMAL, hitNormal receives the face normal of the last step taken (the
// value iterate_dda leaves in it), computed once from the last step mask.
template <bool NORMAL>
__device__ __forceinline__ bool step_map_flat(const DNsceneView& s, RayCtx& c, f3 dir, f3 inv, f3& p,
                                              bool ignoreFirst, DVoxel& vox, f3& colorAdd, float& colorMult,
                                              f3& hitNormal)
{
    uint32_t lastMask = 0xFFFFFFFFu;
    bool requested = false;
    colorAdd = splat(0.0f);
    colorMult = 1.0f;
    uint32_t sMap = 0, sVox = 0;
    c.hitIndex = 0xFFFFFFFFu;
    DDA d;
    init_dda(dir, inv, p, d);
    f3 lsd = splat(0.0f);
    const int32_t X = (int32_t)s.mapX, Y = (int32_t)s.mapY, Z = (int32_t)s.mapZ;
    bool inChunk = false, chunkIgnore = false, hit = false;
    int32_t mx = 0, my = 0, mz = 0;      // map cell while walking a chunk
    f3 mside = splat(0.0f), crp0 = splat(0.0f);
    uint32_t chunkIdx = 0;
    Mask512 m;
    const uint4* rq = nullptr;
    for(;;)
    {
        const int32_t lx = inChunk ? 8 : X, ly = inChunk ? 8 : Y, lz = inChunk ? 8 : Z;
        const bool inb = d.px >= 0 && d.py >= 0 && d.pz >= 0 && d.px < lx && d.py < ly && d.pz < lz;
        if(!inb)
        {
            if(!inChunk) break;                              // left the map
            d.px = mx; d.py = my; d.pz = mz;                 // chunk exhausted: resume the map walk
            d.side = mside;
            inChunk = false;
        }
        else if(!inChunk)
        {
            sMap++;
            const uint32_t idx = (uint32_t)d.px + s.mapX * ((uint32_t)d.py + s.mapY * (uint32_t)d.pz);
            if(occ_bit(s.occLoaded, idx))
            {
                if(s.lastUsed[idx] != 0u) s.lastUsed[idx] = 0u;
                rq = reinterpret_cast<const uint4*>(s.chunks + idx);
                m.q[0] = rq[0]; m.q[1] = rq[1]; m.q[2] = rq[2]; m.q[3] = rq[3];
                f3 upd = p + dir * (minc(lsd) - DN_EPS);
                f3 crp = (upd - mk((float)d.px, (float)d.py, (float)d.pz)) * 8.0f;
                crp0 = vclamp(crp, DN_EPS, 8.0f - DN_EPS);
                mx = d.px; my = d.py; mz = d.pz;
                mside = d.side;
                chunkIdx = idx;
                // init_dda at chunk level: delta and step are unchanged
                f3 f = vfloor(crp0);
                d.px = (int32_t)f.x; d.py = (int32_t)f.y; d.pz = (int32_t)f.z;
                f3 sg = mk((float)d.sx, (float)d.sy, (float)d.sz);
                d.side = ((sg * (f - crp0) + sg * 0.5f) + 0.5f) * d.delta;
                lsd = splat(0.0f);
                chunkIgnore = ignoreFirst;
                inChunk = true;
                continue;                                    // first chunk cell on the next trip
            }
            else if(s.hasUnloaded && !requested && occ_bit(s.occPresent, idx))
            {
                s.flags[idx] = 3u;
                requested = true;
            }
        }
        else
        {
            sVox++;
            const uint32_t li = (uint32_t)(d.px + 8 * (d.py + 8 * d.pz));
            if(mask_bit(m, li) && !chunkIgnore)
            {
                uint4 tail = rq[4];
                uint32_t partial[3] = {tail.x, tail.y, tail.z};
                uint32_t vi = tail.w + rank_in(m, partial, li);
                uint4 cw = s.voxels[vi];
                vox = decompress(cw);
                const DNmat& mat = s.mats[vox.material];
                uint32_t thisID = (cw.y & 0xFFFFFF00u) | (cw.x >> 24);
                if(mat.opacity == 1.0f)
                {
                    f3 crp = crp0 + dir * (minc(lsd) + DN_EPS);
                    p = mk((float)mx, (float)my, (float)mz) + crp * 0.125f;
                    c.hitIndex = vi;
                    hit = true;
                    break;
                }
                else if(c.lastVoxID != thisID)
                {
                    float op = mat.opacity;
                    colorAdd = colorAdd + (vox.albedo * (colorMult * op)) * ld3(s.sunStrength);
                    colorMult = colorMult * (1.0f - op);
                    c.lastVoxID = thisID;
                    c.lastVoxRefract = mat.refractIndex;
                }
            }
            else if(c.lastVoxID != 255u)
            {
                c.lastVoxID = 255u;
                c.lastVoxRefract = 1.0f;
            }
        }
        // common step for both levels: lastSideDist = sideDist; iterate
        lsd = d.side;
        {
            f3 sd = d.side;
            bool bx = sd.x <= gmin(sd.y, sd.z);
            bool by = sd.y <= gmin(sd.z, sd.x);
            bool bz = sd.z <= gmin(sd.x, sd.y);
            d.side.x = sd.x + (bx ? 1.0f : 0.0f) * d.delta.x;
            d.side.y = sd.y + (by ? 1.0f : 0.0f) * d.delta.y;
            d.side.z = sd.z + (bz ? 1.0f : 0.0f) * d.delta.z;
            d.px += bx ? d.sx : 0; d.py += by ? d.sy : 0; d.pz += bz ? d.sz : 0;
            if(NORMAL) lastMask = (bx ? 1u : 0u) | (by ? 2u : 0u) | (bz ? 4u : 0u);
        }
        if(inChunk) chunkIgnore = false;
        else ignoreFirst = false;
    }
    (void)chunkIdx;
    if(NORMAL && lastMask != 0xFFFFFFFFu)
    {
        hitNormal.x = ((lastMask & 1u) ? 1.0f : 0.0f) * (float)(-d.sx);
        hitNormal.y = ((lastMask & 2u) ? 1.0f : 0.0f) * (float)(-d.sy);
        hitNormal.z = ((lastMask & 4u) ? 1.0f : 0.0f) * (float)(-d.sz);
    }
    c.sMap = sMap;
    c.sVox = sVox;
    c.cRays++;
    c.cMap += sMap;
    c.cVox += sVox;
    c.cHits += hit ? 1u : 0u;
    return hit;
}

// texture(samplerCube, dir) with GL_LINEAR filtering, GL_CLAMP_TO_EDGE and
// RGB8 texels, as the reference creates its sky cubemap (main.c:335-353).
// Face selection and (s, t) follow the OpenGL cube map table (major axis
// +X: sc = -z, tc = -y; -X: sc = z, tc = -y; +Y: sc = x, tc = z; -Y: sc = x,
// tc = -z; +Z: sc = x, tc = -y; -Z: sc = -x, tc = -y; s = (sc/|ma| + 1)/2);
// on equal magnitudes x wins over y over z (the specification leaves it to
// the implementation). Texel (i, j) of a face is row j of its upload.
__device__ __forceinline__ f3 cube_texel(const uchar4* face, uint32_t n, int32_t i, int32_t j)
{
    i = i < 0 ? 0 : (i > (int32_t)n - 1 ? (int32_t)n - 1 : i);
    j = j < 0 ? 0 : (j > (int32_t)n - 1 ? (int32_t)n - 1 : j);
    const uchar4 t = face[(uint32_t)j * n + (uint32_t)i];
    return mk((float)t.x / 255.0f, (float)t.y / 255.0f, (float)t.z / 255.0f);
}

__device__ __forceinline__ f3 cube_sample(const uchar4* cube, uint32_t n, f3 d)
{
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    uint32_t f;
    float sc, tc, ma;
    if(ax >= ay && ax >= az) { f = d.x >= 0.0f ? 0u : 1u; sc = d.x >= 0.0f ? -d.z : d.z; tc = -d.y; ma = ax; }
    else if(ay >= az) { f = d.y >= 0.0f ? 2u : 3u; sc = d.x; tc = d.y >= 0.0f ? d.z : -d.z; ma = ay; }
    else { f = d.z >= 0.0f ? 4u : 5u; sc = d.z >= 0.0f ? d.x : -d.x; tc = -d.y; ma = az; }
    const float s = (sc / ma + 1.0f) * 0.5f, t = (tc / ma + 1.0f) * 0.5f;
    const float u = s * (float)n - 0.5f, v = t * (float)n - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    const int32_t i0 = (int32_t)fu, j0 = (int32_t)fv;
    const uchar4* face = cube + (size_t)f * n * n;
    const f3 c00 = cube_texel(face, n, i0, j0), c10 = cube_texel(face, n, i0 + 1, j0);
    const f3 c01 = cube_texel(face, n, i0, j0 + 1), c11 = cube_texel(face, n, i0 + 1, j0 + 1);
    return ((c00 * ((1.0f - a) * (1.0f - b)) + c10 * (a * (1.0f - b))) + c01 * ((1.0f - a) * b)) + c11 * (a * b);
}

// sky_color, voxelShared.comp:289-295
__device__ __forceinline__ f3 sky_color(const DNsceneView& s, f3 dir)
{
    if(s.cube) return cube_sample(s.cube, s.cubeSize, dir);
    return vmix(ld3(s.skyBot), ld3(s.skyTop), dir.y * 0.5f + 1.0f);
}

// wave-wide sum and one atomic per wave into a u64 counter
__device__ __forceinline__ void wave_add_u64(uint64_t* dst, uint32_t v)
{
    uint64_t x = v;
    for(int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if((threadIdx.x & 63u) == 0 && x) atomicAdd((unsigned long long*)dst, (unsigned long long)x);
}

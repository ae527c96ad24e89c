// Sphere carve/fill edits on the device (voxelShapes.c:239-333 applied to the
// words of voxel.c's CPU chunks). One block per touched map tile holds the
// tile's 512 voxel words in registers and applies, in edit order, every sphere
// of the batch that can touch the tile. The host replays the resulting chunk
// add/remove events in the reference's order (edit by edit, tiles in z, y, x
// order), so chunk slots and dirty order come out as with sequential edits.
#include "dn_internal.h"

namespace {

constexpr uint32_t EMPTY = 255u;

// DN_compress_voxel's normal quantisation (voxel.c:1293-1297) with quickmath's
// QM_MIN/QM_MAX (x < y ? x : y): a NaN component (fromCenter / 0 at a sphere's
// centre voxel) quantises as 1.
__device__ __forceinline__ uint32_t quant(float x)
{
    x = x < 1.0f ? x : 1.0f;
    x = x > -1.0f ? x : -1.0f;
    return (uint32_t)(((int)(x * 255.0f) + 255) / 2);
}

__device__ __forceinline__ uint32_t normal_word(uint32_t material, float nx, float ny, float nz)
{
    return (material << 24) | (quant(nx) << 16) | (quant(ny) << 8) | quant(nz);
}

// One voxel's part of one sphere pass. `present` is the tile's chunk-present
// flag at the start of the pass (the reference reads it once per tile).
// Returns whether the voxel is written; `w` receives the new words, `delta`
// the change of the tile's stored-voxel count.
__device__ __forceinline__ bool sphere_voxel(const DNeditSphere& e, bool present, int gx, int gy, int gz, uint2 old,
                                             uint2& w, int& delta)
{
    delta = 0;
    const float fx = (float)gx - e.c[0], fy = (float)gy - e.c[1], fz = (float)gz - e.c[2];
    const float d2 = fx * fx + fy * fy + fz * fz;
    const bool solid = (old.x >> 24) != EMPTY;
    if(e.material != EMPTY)
    {
        if(!(d2 < e.r2) || (present && solid)) return false;
        const float m = fmaxf(fmaxf(fabsf(fx), fabsf(fy)), fabsf(fz));
        const float s = 1.0f / m;
        w = make_uint2(normal_word(e.material, fx * s, fy * s, fz * s), e.albedo);
        delta = solid ? 0 : 1;
        return true;
    }
    if(!present) return false;                 // empty into a missing chunk: nothing
    if(d2 < e.r2)
    {
        w = make_uint2(e.carveNormal, e.albedo);
        delta = solid ? -1 : 0;
        return true;
    }
    if(e.flip && d2 < e.r12 && solid)
    {
        const float m = fmaxf(fmaxf(fabsf(fx), fabsf(fy)), fabsf(fz));
        const float s = -1.0f / m;
        // DN_get_voxel then DN_set_voxel: material kept, albedo keeps its colour bytes
        w = make_uint2(normal_word(old.x >> 24, fx * s, fy * s, fz * s), old.y & 0xFFFFFF00u);
        return true;
    }
    return false;
}

// 256 threads; thread t owns voxels t and t + 256 of the tile, voxel (x, y, z)
// at x*64 + y*8 + z (the CPU chunk's [x][y][z] order).
__global__ __launch_bounds__(256) void k_sphere_edits(DNeditParams P)
{
    __shared__ int red[2][4][2];
    const uint32_t u = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t t = P.tileIds[u];
    const int tx = (int)(t % P.mapX), ty = (int)((t / P.mapX) % P.mapY), tz = (int)(t / (P.mapX * P.mapY));
    uint32_t info = P.tileInfo[u];
    bool present = (info >> 31) != 0u;
    int count = (int)(info & 0x7FFFFFFFu);
    uint2* words = P.words + (size_t)t * 512u;
    const uint2 cleared = make_uint2(0xFFFFFFFFu, 0u);
    uint2 v[2];
    int gx[2], gy[2], gz[2];
    for(int k = 0; k < 2; k++)
    {
        const uint32_t i = tid + 256u * (uint32_t)k;
        v[k] = present ? words[i] : cleared;
        gx[k] = tx * 8 + (int)(i >> 6);
        gy[k] = ty * 8 + (int)((i >> 3) & 7u);
        gz[k] = tz * 8 + (int)(i & 7u);
    }
    bool modified = false;
    for(uint32_t q = P.pairStart[u], qe = P.pairStart[u + 1]; q < qe; q++)
    {
        const DNeditSphere e = P.spheres[P.pairSphere[q]];
        uint2 w[2];
        int d[2];
        bool wr[2];
        for(int k = 0; k < 2; k++) wr[k] = sphere_voxel(e, present, gx[k], gy[k], gz[k], v[k], w[k], d[k]);
        // block sums of the count change and of the writes
        const int inc = __popcll(__ballot(d[0] > 0)) + __popcll(__ballot(d[1] > 0));
        const int dec = __popcll(__ballot(d[0] < 0)) + __popcll(__ballot(d[1] < 0));
        const int any = __popcll(__ballot(wr[0] || wr[1]));
        const int buf = (int)(q & 1u);         // alternate buffers: one barrier per pass
        if(lane == 0) { red[buf][wave][0] = inc - dec; red[buf][wave][1] = any; }
        __syncthreads();
        const int delta = red[buf][0][0] + red[buf][1][0] + red[buf][2][0] + red[buf][3][0];
        const bool wrote = (red[buf][0][1] | red[buf][1][1] | red[buf][2][1] | red[buf][3][1]) != 0;
        uint32_t ev = 0;
        if(wrote)
        {
            ev |= DN_EDIT_MODIFIED;
            modified = true;
            for(int k = 0; k < 2; k++)
                if(wr[k]) v[k] = w[k];
            const int before = count;
            count += delta;
            if(e.material != EMPTY && !present)
            {
                ev |= DN_EDIT_ADDED;            // the first solid voxel adds the chunk
                present = true;
            }
            else if(e.material == EMPTY && before > 0 && count <= 0)
            {
                ev |= DN_EDIT_REMOVED;          // the last solid voxel removes (and clears) it
                present = false;
                count = 0;
                v[0] = v[1] = cleared;
            }
        }
        if(tid == 0) P.events[P.pairSeq[q]] = (uint8_t)ev;
    }
    for(int k = 0; k < 2; k++) words[tid + 256u * (uint32_t)k] = v[k];
    if(tid == 0) P.tileOut[u] = (present ? 0x80000000u : 0u) | (uint32_t)count | (modified ? 0x40000000u : 0u);
}

// Copies the final words of the listed tiles into a contiguous buffer for the
// host (one block per tile).
__global__ __launch_bounds__(256) void k_edit_gather(const uint2* __restrict__ words, const uint32_t* tiles,
                                                     uint2* out)
{
    const uint2* src = words + (size_t)tiles[blockIdx.x] * 512u;
    uint2* dst = out + (size_t)blockIdx.x * 512u;
    dst[threadIdx.x] = src[threadIdx.x];
    dst[threadIdx.x + 256u] = src[threadIdx.x + 256u];
}

// Scatters staged host words of the listed tiles into the device copy.
__global__ __launch_bounds__(256) void k_edit_scatter(uint2* words, const uint32_t* tiles, const uint2* __restrict__ in)
{
    uint2* dst = words + (size_t)tiles[blockIdx.x] * 512u;
    const uint2* src = in + (size_t)blockIdx.x * 512u;
    dst[threadIdx.x] = src[threadIdx.x];
    dst[threadIdx.x + 256u] = src[threadIdx.x + 256u];
}

}  // namespace

int dn_launch_edit_scatter(uint2* words, const uint32_t* tiles, const uint2* in, uint32_t n, hipStream_t stream)
{
    if(n) hipLaunchKernelGGL(k_edit_scatter, dim3(n), dim3(256), 0, stream, words, tiles, in);
    return (int)hipGetLastError();
}

int dn_launch_sphere_edits(const DNeditParams& p, uint32_t numTiles, hipStream_t stream)
{
    if(numTiles) hipLaunchKernelGGL(k_sphere_edits, dim3(numTiles), dim3(256), 0, stream, p);
    return (int)hipGetLastError();
}

int dn_launch_edit_gather(const uint2* words, const uint32_t* tiles, uint2* out, uint32_t n, hipStream_t stream)
{
    if(n) hipLaunchKernelGGL(k_edit_gather, dim3(n), dim3(256), 0, stream, words, tiles, out);
    return (int)hipGetLastError();
}

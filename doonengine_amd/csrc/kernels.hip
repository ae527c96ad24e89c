// HIP kernels for gfx950: draw cast, lighting pass, visible-chunk compaction,
// chunk upload scatter, ray batches. Launchers at the end.
#include "dn_device.hpp"

namespace {

// ---------------------------------------------------------------------------
// draw cast (assets/shaders/voxelDraw.comp:24-147). A 256-thread block covers a
// 16x16 pixel tile; each wave64 covers an 8x8 quadrant of it.

struct PixelHit { uint32_t voxelIndex, mapSteps, voxelSteps, pad; float face[3]; };

__device__ f3 voxel_color(const DNsceneView& s, uint32_t viewMode, DVoxel vox, f3 ca, float cm, f3 hitNormal)
{
    const DNmat& m = s.mats[vox.material];
    vox.spec = vox.spec * m.specular;
    vox.diffuse = vox.diffuse * (1.0f - m.specular);
    switch(viewMode)
    {
    case 0: {
        f3 solid = m.emissive ? vox.albedo * emission_of(m) : vox.diffuse * vox.albedo + vox.spec;
        return solid * cm + ca;
    }
    case 1: return vox.albedo;
    case 2: return m.emissive ? vox.albedo : vox.diffuse;
    case 3: return m.emissive ? vox.albedo : vox.spec;
    case 4: return vabs(vox.normal);
    default: return vabs(hitNormal);
    }
}

// One pixel's primary ray (voxelDraw.comp:24-147). The per-ray refraction state
// starts afresh for every pixel, as the shader's per-invocation globals do.
__device__ __forceinline__ void draw_pixel(const DNdrawParams& P, const DNsceneView& s, RayCtx& c, uint32_t x,
                                           uint32_t y)
{
    c.lastVoxID = 255u; c.lastVoxRefract = 1.0f;
    const float* M = P.invViewProj;
    float sx = (float)x / (float)P.width * 2.0f - 1.0f;
    float sy = (float)y / (float)P.height * 2.0f - 1.0f;
    f3 org = mk(P.invView[12], P.invView[13], P.invView[14]);
    f3 p = org;
    f3 q = mk(M[0] * sx + M[4] * sy + M[8] * 0.0f + M[12],
              M[1] * sx + M[5] * sy + M[9] * 0.0f + M[13],
              M[2] * sx + M[6] * sy + M[10] * 0.0f + M[14]);
    f3 dir = vnorm(q) + DN_EPS;
    f3 inv = vinv(dir);
    f3 ms = mk((float)s.mapX, (float)s.mapY, (float)s.mapZ);
    f3 color;
    float depth = -1.0f;
    PixelHit info = {0xFFFFFFFFu, 0u, 0u, 0u, {0.0f, 0.0f, 0.0f}};
    f3 tMin = (splat(0.0f) - p) * inv, tMax = (ms - p) * inv;
    f3 t1 = vmin(tMin, tMax), t2 = vmax(tMin, tMax);
    float tNear = gmax(gmax(t1.x, t1.y), t1.z), tFar = gmin(gmin(t2.x, t2.y), t2.z);
    if(tNear > tFar || tFar < 0.0f)
        color = vpow(sky_color(s, dir), 2.2f);
    else
    {
        if(tNear > 0.0f) p = p + dir * (tNear + DN_EPS);
        // normal_AABB, voxelShared.comp:276-284
        f3 cen = (splat(0.0f) + ms) * 0.5f;
        f3 pc = p - cen;
        f3 dd = (ms - splat(0.0f)) * 0.5f;
        float bias = 1.0f + DN_EPS;
        f3 tr = mk(truncf(pc.x / dd.x * bias), truncf(pc.y / dd.y * bias), truncf(pc.z / dd.z * bias));
        f3 nrm = vnorm(tr);
        DVoxel vox = {};
        f3 ca; float cm;
        bool hit = step_map<true, true>(s, c, dir, inv, p, false, nrm, vox, ca, cm);
        info.mapSteps = c.sMap; info.voxelSteps = c.sVox;
        if(hit)
        {
            f3 fp = vfloor(p);
            if(fp.x >= 0.0f && fp.y >= 0.0f && fp.z >= 0.0f && fp.x < ms.x && fp.y < ms.y && fp.z < ms.z)
            {
                uint32_t idx = (uint32_t)fp.x + s.mapX * ((uint32_t)fp.y + s.mapY * (uint32_t)fp.z);
                if(!(s.flags[idx] & 4u)) atomicOr(&s.flags[idx], 4u);
            }
            f3 dv = p - org;
            depth = sqrtf(dot3(dv, dv));
            color = voxel_color(s, P.viewMode, vox, ca, cm, nrm);
            info.voxelIndex = c.hitIndex;
        }
        else
            color = vpow(sky_color(s, dir), 2.2f) * cm + ca;
        info.face[0] = nrm.x; info.face[1] = nrm.y; info.face[2] = nrm.z;
    }
    color = vpow(color, 0.4545f);
    size_t o = (size_t)y * P.width + x;
    P.out[o] = make_float4(color.x, color.y, color.z, depth);
    if(P.hitOut) reinterpret_cast<PixelHit*>(P.hitOut)[o] = info;
}

// Persistent draw: each wave takes 8x8 pixel quads (numbered row by row) from
// a device counter, lane l drawing pixel (l & 7, l >> 3) of the quad. Chunk
// masks go to per-lane LDS rows on chunk entry; the chunk-grid occupancy bits
// are read through L2 (staging the bitmap in LDS per block measured 0.75 vs
// 0.71 ms at 1920x1080: neighbouring primary rays read the same words). The
// last wave to finish resets the quad counter for the next launch.
#ifndef DN_DRAW_WAVES
#define DN_DRAW_WAVES 4
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DN_DRAW_WAVES, DN_DRAW_WAVES)))
void k_draw(DNdrawParams P)
{
    const uint32_t lane = threadIdx.x & 63u;
    __shared__ uint32_t maskRows[DN_MASK_LDS_WORDS(256)];
    const DNsceneView sv = P.scene;
    RayCtx c;
    ctx_init(c);
    c.maskRow = maskRows + threadIdx.x * DN_MASK_ROW;
    const uint32_t qx = (P.width + 7u) / 8u, nq = qx * ((P.height + 7u) / 8u);
    uint32_t drawn = 0;
    // The first 3/4 of the quads are dealt out statically (wave w takes quads
    // w, w + W, w + 2W, ...) and the rest from the device counter, in order,
    // to even out the waves' ends. Every quad from the counter costs one
    // atomic on that single word, which all 8 XCDs share; at one atomic per
    // quad those serialise to a floor of about 0.5 ms per 1920x1080 frame
    // (0.70 ms measured against 0.52 ms with the static share).
    const uint32_t waves = gridDim.x * 4u, wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t perWave = (nq / 4u * 3u) / waves, head = perWave * waves;
    for(uint32_t k = 0;; k++)
    {
        uint32_t q = wave + k * waves;
        if(k >= perWave)
        {
            if(lane == 0) q = atomicAdd(P.quadCounter, 1u);
            q = __shfl(q, 0, 64) + head;
        }
        if(q >= nq) break;
        const uint32_t x = (q % qx) * 8u + (lane & 7u), y = (q / qx) * 8u + (lane >> 3);
        if(x < P.width && y < P.height)
        {
            drawn++;
            draw_pixel(P, sv, c, x, y);
        }
    }
    if(P.counters)
    {
        wave_add_u64(P.counters + CNT_RAYS, c.cRays);
        wave_add_u64(P.counters + CNT_MAPSTEPS, c.cMap);
        wave_add_u64(P.counters + CNT_VOXSTEPS, c.cVox);
        wave_add_u64(P.counters + CNT_HITS, c.cHits);
        wave_add_u64(P.counters + CNT_PIXELS, drawn);
    }
    if(lane == 0)
    {
        // every wave has taken its last quad before it counts itself out
        __threadfence();
        const uint32_t done = atomicAdd(P.quadCounter + 1, 1u);
        if(done == gridDim.x * 4u - 1u)
        {
            atomicExch(P.quadCounter, 0u);
            atomicExch(P.quadCounter + 1, 0u);
        }
    }
}

}  // namespace

#include "dn_light.hpp"

namespace {

// ---------------------------------------------------------------------------
// lighting pass: one lane per stored voxel; a wave takes two consecutive
// 32-voxel requests (normally the two halves of one chunk's voxel range) from
// a device work counter.

// Waves per SIMD the lighting kernel is compiled for. At 4 the register
// allocator keeps 128 VGPRs (a few spills, outside the DDA loops); measured
// 12% faster than the 154 VGPRs / 3 waves it would take unconstrained.
#ifndef DN_LIGHT_WAVES
#define DN_LIGHT_WAVES 4
#endif
// Largest chunk-grid occupancy bitmap staged in LDS (bytes): 64 KiB covers
// 524 288 tiles; bigger maps read it from global memory.
#define DN_OCC_LDS_MAX 65536u

template <bool LDS_OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DN_LIGHT_WAVES, DN_LIGHT_WAVES)))
void k_lighting(DNlightParams P)
{
    const uint32_t lane = threadIdx.x & 63u;
    __shared__ uint32_t maskRows[DN_MASK_LDS_WORDS(256)];
    RayCtx c;
    ctx_init(c);
    c.maskRow = lane_mask_row(maskRows);
    uint32_t lit = 0;
    const uint32_t nreq = *P.numRequests;
    // the chunk-grid occupancy bits, staged in LDS: every chunk-grid trip of
    // every ray tests one of them
    extern __shared__ uint32_t occLds[];
    DNsceneView sv = P.scene;
    if(LDS_OCC)
    {
        const uint32_t words = (P.tiles + 31u) / 32u;
        for(uint32_t i = threadIdx.x; i < words; i += 256u) occLds[i] = P.scene.occLoaded[i];
        __syncthreads();
        sv.occLoaded = occLds;
    }
    for(;;)
    {
        uint32_t base = 0;
        if(lane == 0) base = atomicAdd(P.workCounter, 2u);
        base = __shfl(base, 0, 64);
        if(base >= nreq) break;
        const uint32_t r = base + (lane >> 5);
        if(r < nreq) light_voxel(P, sv, r, lane & 31u, c, lit, false);
    }
    if(P.counters)
    {
        wave_add_u64(P.counters + CNT_RAYS, c.cRays);
        wave_add_u64(P.counters + CNT_MAPSTEPS, c.cMap);
        wave_add_u64(P.counters + CNT_VOXSTEPS, c.cVox);
        wave_add_u64(P.counters + CNT_HITS, c.cHits);
        wave_add_u64(P.counters + CNT_LIT, lit);
    }
}

// Direct rays of the pass's voxels, traced ahead of k_lighting. The N direct
// directions are the same for every voxel and lie in a narrow cone around the
// sun, and a voxel's N rays share their start point: their lengths are alike.
// So a wave holds 64 / N voxels x N directions (lane = voxel * N + direction)
// instead of 64 voxels x one direction, and the longest ray of a wave is
// rarely much longer than the rest. Each voxel's terms are then summed in
// direction order on the lane of its first direction, exactly as shadow_ray
// sums them, and stored per request slot (w = 0). Every ray here starts with
// the shader's lastVoxID = 255; if one of a voxel's rays ends with another
// value (it stopped inside a transparent voxel), the next ray of the
// sequential loop would have started differently, so the slot is marked
// (w = 1) and the pass traces that voxel's rays itself. The rays' counters go
// with the term (the pass adds them whenever it uses it), so the pass's counts
// are those of tracing in sequence. Voxels without a diffuse term (specular
// >= 1) are skipped. A wave takes one request (32 slots) at a time.
// Waves per SIMD the direct-ray kernel is compiled for: 96 VGPRs, no spills.
// Measured on the full-frame workload (lighting pass) with the sums spread
// over 8 lanes per voxel: 4 waves 4.73 ms, 5 waves 4.38, 6 waves (80 VGPRs,
// a few spills) 4.67, 7 waves 4.67. With one summing lane per voxel 6 waves
// had been best (4 waves 6.55 ms, 5 waves 5.88, 6 waves 5.49, 7 waves 5.93).
// The chunk-grid occupancy bitmap is read from L2 here: staging it in LDS
// (16 KiB per block on that map) would hold the kernel to 4 waves.
#ifndef DN_DIRECT_WAVES
#define DN_DIRECT_WAVES 5
#endif
// The batched direct pass (384 uneven volumes, 16 directions: 4 voxels per
// wave) measured 49.0 ms of graph time at 6 waves against 52.0 at 5 (with the
// wave's origin row: 47.7 against 49.8).
#ifndef DN_DIRECT_BATCH_WAVES
#define DN_DIRECT_BATCH_WAVES 6
#endif
// The direct rays of one request's 32 slots: lane = voxel * N + direction,
// 64 / N voxels at a time. `row` is the lane's mask row, `voxRows` the row of
// its voxel's first direction.
__device__ __forceinline__ void direct_request(const DNlightParams& P, const DNsceneView& sv, uint32_t r, RayCtx& c,
                                               uint32_t* row, const uint32_t* voxRows, uint32_t* originRow,
                                               uint32_t N, uint32_t per, uint32_t v, uint32_t k, f3 dir, f3 inv,
                                               f3 sun)
{
    const uint32_t req = P.requests[r];
    const DNchunkRec* rec = sv.chunks + (req >> 4);
    // every ray of the request starts in the request's chunk: its mask is
    // staged once, in the wave's origin row, for all of them
    const uint32_t lane = threadIdx.x & 63u;
    __builtin_amdgcn_wave_barrier();
    if(lane < 16u) originRow[lane] = rec->mask[lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    c.sharedRow = originRow;
    c.sharedMap = req >> 4;
    // each slot's start point once per request: lane l prepares slot l & 31,
    // the tracing lanes fetch theirs by shuffle
    const uint32_t myNum = (threadIdx.x & 31u) + (req & 15u) * 32u;
    int32_t selL = select_in(rec->mask, myNum);
    f3 pL = splat(0.0f);
    if(selL >= 0)
    {
        const DVoxel me = decompress(sv.voxels[rec->voxelIndex + myNum]);
        if(sv.mats[me.material].specular < 1.0f) pL = voxel_origin(rec, selL, me.normal);
        else selL = -1;                                              // no diffuse term
    }
    for(uint32_t g = 0; g < 32u; g += per)
    {
        const uint32_t lane32 = g + v;
        const bool mine = v < per && lane32 < 32u;
        const int32_t src = mine ? (int32_t)lane32 : 0;
        int32_t sel = __shfl(selL, src, 64);
        f3 pp = mk(__shfl(pL.x, src, 64), __shfl(pL.y, src, 64), __shfl(pL.z, src, 64));
        if(!mine) sel = -1;
        float4 term = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const uint32_t r0 = c.cRays, m0 = c.cMap, v0 = c.cVox, h0 = c.cHits;
        if(sel >= 0)
        {
            c.lastVoxID = 255u; c.lastVoxRefract = 1.0f;
            DVoxel hv = {};
            f3 ca;
            float cm;
            if(!light_trace_uniform(sv, c, dir, inv, pp, hv, ca, cm))
            {
                const f3 t = sun * cm + ca;
                term = make_float4(t.x, t.y, t.z, 1.0f);
            }
            if(c.lastVoxID != 255u) term.w += 2.0f;    // ended inside a transparent voxel
        }
        __builtin_amdgcn_wave_barrier();      // every lane's walk is done with its row
        // the row: the term's x, y, z (0 when the ray does not reach the sun),
        // whether the ray ended inside a transparent voxel, the ray's counters
        row[0] = __float_as_uint(term.x); row[1] = __float_as_uint(term.y); row[2] = __float_as_uint(term.z);
        row[3] = term.w >= 2.0f ? 1u : 0u;
        row[4] = c.cRays - r0; row[5] = c.cMap - m0; row[6] = c.cVox - v0; row[7] = c.cHits - h0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Lane k of a voxel sums word k of the voxel's N rows (a lane takes
        // words k, k + N, ... when N < 8), which is word k of its DNdirectTerm:
        // x, y, z in direction order as shadow_ray adds them (a term that
        // does not reach the sun is +0 and leaves the sum as it is; every sum
        // is >= +0), divided by N; the transparent-end flags (any); the
        // counters (integers, any order).
        if(sel >= 0)
            for(uint32_t wd = k; wd < 8u; wd += N)
            {
                float fs = 0.0f;
                uint32_t us = 0u;
                const uint32_t* q = voxRows + wd;
                for(uint32_t j = 0; j < N; j++, q += DN_MASK_ROW)
                {
                    const uint32_t x = *q;
                    fs = fs + __uint_as_float(x);
                    us += x;
                }
                const size_t slot = (size_t)r * 32u + lane32;
                if(slot < P.directSlots)
                    reinterpret_cast<uint32_t*>(P.directOut + slot)[wd] =
                        wd < 3u ? __float_as_uint(fs / (float)N) : (wd == 3u ? __float_as_uint(us ? 1.0f : 0.0f) : us);
            }
        __builtin_amdgcn_wave_barrier();
    }
    c.sharedMap = 0xFFFFFFFFu;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DN_DIRECT_WAVES, DN_DIRECT_WAVES)))
void k_direct(DNlightParams P)
{
    const uint32_t lane = threadIdx.x & 63u;
    // per lane: the chunk-mask row of its walks, then (between walks) its
    // ray's term and counters for the summing lane of its voxel
    __shared__ uint32_t maskRows[DN_MASK_LDS_WORDS(256)];
    __shared__ uint32_t originRows[4 * DN_MASK_ROW];     // per wave: the request chunk's mask
    const DNsceneView sv = P.scene;
    const uint32_t N = P.numDirectRays;
    RayCtx c;
    ctx_init(c);
    uint32_t* row = maskRows + threadIdx.x * DN_MASK_ROW;
    c.maskRow = row;
    const uint32_t nreq = *P.numRequests;
    const uint32_t per = 64u / N, v = lane / N, k = lane - v * N;
    const f3 dir = ld3(P.directRays[k]), inv = ld3(P.directRays[k] + 3), sun = ld3(sv.sunStrength);
    // row of this voxel's first direction (lanes past the last whole voxel get
    // their wave's first row; they never sum)
    const uint32_t* voxRows = maskRows + (v < per ? threadIdx.x - k : threadIdx.x - lane) * DN_MASK_ROW;
    for(;;)
    {
        uint32_t r = 0;
        if(lane == 0) r = atomicAdd(P.workCounter + 4, 1u);
        r = __shfl(r, 0, 64);
        if(r >= nreq) break;
        direct_request(P, sv, r, c, row, voxRows, originRows + (threadIdx.x >> 6) * DN_MASK_ROW, N, per, v, k, dir,
                       inv, sun);
    }
}

// k_direct over the volumes of a batch, with the work units, queues and
// per-queue counters of k_lighting_batch (its own counters). Volumes whose pass
// traces its direct rays itself (no directOut) are skipped.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DN_DIRECT_BATCH_WAVES, DN_DIRECT_BATCH_WAVES)))
void k_direct_batch(DNbatchParams B, const DNlightParams* __restrict__ scenes)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t home = blockIdx.x % B.queues;
    __shared__ uint32_t maskRows[DN_MASK_LDS_WORDS(256)];
    __shared__ uint32_t originRows[4 * DN_MASK_ROW];     // per wave: the request chunk's mask
    RayCtx c;
    ctx_init(c);
    uint32_t* row = maskRows + threadIdx.x * DN_MASK_ROW;
    c.maskRow = row;
    for(uint32_t qi = 0; qi < B.queues; qi++)
    {
        const uint32_t q = (home + qi) % B.queues;
        const uint32_t len = B.queueLen[q];
        if(len == 0) continue;
        const uint32_t* pre = B.queuePrefix + (size_t)q * (B.numScenes + 1u);
        const uint32_t total = pre[len];
        for(;;)
        {
            uint32_t u = 0;
            if(lane == 0) u = atomicAdd(&B.directCounters[q], 1u);
            u = __builtin_amdgcn_readfirstlane(__shfl(u, 0, 64));
            if(u >= total) break;
            uint32_t lo = 0, hi = len - 1u;           // last volume with pre[i] <= u
            while(lo < hi)
            {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if(pre[mid] <= u) lo = mid;
                else hi = mid - 1u;
            }
            const uint32_t sid = __builtin_amdgcn_readfirstlane(B.queueScenes[(size_t)q * B.numScenes + lo]);
            const DNlightParams& P = scenes[sid];
            if(!P.directOut) continue;
            // The direct terms are per-frame scratch of the pass, written in a dry
            // run too (the dry lighting pass reads them); volume state is not.
            DNsceneView sv = P.scene;
            if(B.dry)
            {
                sv.hasUnloaded = 0u;
                sv.lastUsed = B.dryLastUsed;
            }
            const uint32_t N = P.numDirectRays;
            const uint32_t per = 64u / N, v = lane / N, k = lane - v * N;
            const f3 dir = ld3(P.directRays[k]), inv = ld3(P.directRays[k] + 3), sun = ld3(sv.sunStrength);
            const uint32_t* voxRows = maskRows + (v < per ? threadIdx.x - k : threadIdx.x - lane) * DN_MASK_ROW;
            const uint32_t nreq = *P.numRequests;
            const uint32_t r0 = (u - pre[lo]) * B.unit;
            for(uint32_t j = 0; j < B.unit && r0 + j < nreq; j++)
                direct_request(P, sv, r0 + j, c, row, voxRows, originRows + (threadIdx.x >> 6) * DN_MASK_ROW, N,
                               per, v, k, dir, inv, sun);
        }
    }
}

// Pass epilogue over one 32-tile word of the bitmaps: clear the visible bit and
// bump the sample count of lit chunks, then set the visible bit of chunks
// reached by specular rays from visible chunks.
__device__ __forceinline__ void epilogue_word(uint32_t* flags, DNchunkRec* chunks, uint32_t* litBits,
                                              uint32_t* visProp, uint32_t w)
{
    uint32_t lit = litBits[w], prop = visProp[w];
    if(!(lit | prop)) return;
    uint32_t any = lit | prop;
    while(any)
    {
        uint32_t b = (uint32_t)__builtin_ctz(any);
        any &= any - 1u;
        uint32_t t = w * 32u + b;
        uint32_t f = flags[t];
        if(lit & (1u << b)) { f &= ~4u; chunks[t].samples += 1u; }
        if(prop & (1u << b)) f |= 4u;
        flags[t] = f;
    }
    litBits[w] = 0u;
    visProp[w] = 0u;
}

// The epilogue of one pass; also resets the pass's work counter.
__global__ __launch_bounds__(256) void k_light_epilogue(uint32_t* flags, DNchunkRec* chunks, uint32_t* litBits,
                                                        uint32_t* visProp, uint32_t words, uint32_t* workCounter)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if(w == 0) { workCounter[0] = 0u; workCounter[4] = 0u; }
    if(w < words) epilogue_word(flags, chunks, litBits, visProp, w);
}

// Copies the lit voxels of a pass from the output pool back into the pool
// (only used when the pass wrote to a separate buffer); grid-stride over the
// n * 32 voxel slots of the requests.
__device__ __forceinline__ void copy_back_range(const uint32_t* requests, uint32_t n, const DNchunkRec* chunks,
                                                const uint4* src, uint4* dst, uint32_t first, uint32_t stride)
{
    for(uint32_t g = first; g < n * 32u; g += stride)
    {
        uint32_t req = requests[g >> 5];
        uint32_t voxNum = (g & 31u) + (req & 15u) * 32u;
        const DNchunkRec* rec = chunks + (req >> 4);
        if(voxNum < rec->numVoxels) dst[rec->voxelIndex + voxNum] = src[rec->voxelIndex + voxNum];
    }
}

__global__ __launch_bounds__(256) void k_copy_back(const uint32_t* requests, const uint32_t* numRequests,
                                                   const DNchunkRec* chunks, const uint4* src, uint4* dst)
{
    copy_back_range(requests, *numRequests, chunks, src, dst, blockIdx.x * blockDim.x + threadIdx.x,
                    gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// visible-chunk compaction (voxel.c:719-762 + 1463-1489). Tiles are visited in
// map-index order (z, y, x); a tile emits ceil(numVoxels/32) request words.

__device__ __forceinline__ uint32_t tile_words(const DNvisibleParams& P, uint32_t t, bool isFresh)
{
    uint32_t f = P.flags[t];
    if(P.split == 0 || (f & 3u) != 2u || !(f & 4u)) return 0u;
    if(t % P.split != P.frameNum && !isFresh) return 0u;
    return (P.chunks[t].numVoxels + 31u) >> 5;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const uint32_t lane = threadIdx.x & 63u;
    for(uint32_t off = 1; off < 64; off <<= 1)
    {
        uint32_t u = __shfl_up(v, off, 64);
        if(lane >= off) v += u;
    }
    return v;
}

__global__ __launch_bounds__(1024) void k_visible_count(DNvisibleParams P)
{
    __shared__ uint32_t waveSums[16];
    const uint32_t t = blockIdx.x * 1024u + threadIdx.x;
    uint32_t cnt = 0;
    if(t < P.tiles)
    {
        bool isFresh = (P.fresh[t >> 5] >> (t & 31u)) & 1u;
        if(!isFresh) P.lastUsed[t] += 1u;
        cnt = tile_words(P, t, isFresh);
    }
    uint32_t s = wave_incl_scan(cnt);
    if((threadIdx.x & 63u) == 63u) waveSums[threadIdx.x >> 6] = s;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        uint32_t tot = 0;
        for(int i = 0; i < 16; i++) tot += waveSums[i];
        P.blockSums[blockIdx.x] = tot;
    }
}

__global__ __launch_bounds__(1024) void k_visible_write(DNvisibleParams P)
{
    __shared__ uint32_t waveSums[16];
    __shared__ uint32_t blockBase;
    const uint32_t t = blockIdx.x * 1024u + threadIdx.x;
    if(threadIdx.x < 64)
    {
        uint32_t acc = 0;
        for(uint32_t b = threadIdx.x; b < blockIdx.x; b += 64) acc += P.blockSums[b];
        for(int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if(threadIdx.x == 0) blockBase = acc;
    }
    uint32_t freshWord = 0;
    uint32_t cnt = 0;
    if(t < P.tiles)
    {
        freshWord = P.fresh[t >> 5];
        cnt = tile_words(P, t, (freshWord >> (t & 31u)) & 1u);
    }
    uint32_t incl = wave_incl_scan(cnt);
    if((threadIdx.x & 63u) == 63u) waveSums[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t waveBase = 0;
    for(uint32_t i = 0; i < (threadIdx.x >> 6); i++) waveBase += waveSums[i];
    uint32_t first = blockBase + waveBase + incl - cnt;
    for(uint32_t b = 0; b < cnt; b++) P.requests[first + b] = (t << 4) | b;
    if(blockIdx.x == gridDim.x - 1 && threadIdx.x == 1023)
        *P.numRequests = first + cnt;
    __syncthreads();
    if(t < P.tiles && (t & 31u) == 0 && freshWord) P.fresh[t >> 5] = 0u;
}

// ---------------------------------------------------------------------------
// upload scatter: one block per operation

__global__ __launch_bounds__(256) void k_upload(uint32_t* flags, uint32_t* lastUsed, DNchunkRec* chunks, uint4* voxels,
                                                uint32_t* fresh, const DNuploadOp* ops, const DNchunkRec* recs,
                                                const uint4* staged)
{
    const DNuploadOp op = ops[blockIdx.x];
    if(op.mode == 2u)
    {
        const DNchunkRec* r = recs + op.recSlot;
        if(threadIdx.x < 8)
            reinterpret_cast<uint4*>(chunks + op.tile)[threadIdx.x] = reinterpret_cast<const uint4*>(r)[threadIdx.x];
        const uint32_t n = r->numVoxels, base = r->voxelIndex;
        for(uint32_t i = threadIdx.x; i < n; i += blockDim.x) voxels[base + i] = staged[op.voxOffset + i];
        if(threadIdx.x == 0)
        {
            flags[op.tile] = 2u | (flags[op.tile] & 4u);
            lastUsed[op.tile] = 0u;
            atomicOr(&fresh[op.tile >> 5], 1u << (op.tile & 31u));
        }
    }
    else if(threadIdx.x == 0)
        flags[op.tile] = op.mode;
}

// ---------------------------------------------------------------------------
// ray batch: one step_map per ray

struct RayHit { uint32_t hit, voxelIndex, mapSteps, voxelSteps; float pos[3]; float normal[3]; };

__global__ __launch_bounds__(256) void k_rays(DNrayParams P)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    RayCtx c;
    ctx_init(c);
    if(i < P.n)
    {
        f3 dir = ld3(P.dirs + 3 * (size_t)i), p = ld3(P.origins + 3 * (size_t)i), nrm = splat(0.0f), ca;
        float cm;
        DVoxel vox = {};
        bool h = step_map_flat<true>(P.scene, c, dir, vinv(dir), p, P.ignoreFirst != 0, vox, ca, cm, nrm);
        RayHit o;
        o.hit = h ? 1u : 0u;
        o.voxelIndex = h ? c.hitIndex : 0xFFFFFFFFu;
        o.mapSteps = c.sMap; o.voxelSteps = c.sVox;
        o.pos[0] = p.x; o.pos[1] = p.y; o.pos[2] = p.z;
        o.normal[0] = nrm.x; o.normal[1] = nrm.y; o.normal[2] = nrm.z;
        reinterpret_cast<RayHit*>(P.out)[i] = o;
    }
    if(P.counters)
    {
        wave_add_u64(P.counters + CNT_RAYS, c.cRays);
        wave_add_u64(P.counters + CNT_MAPSTEPS, c.cMap);
        wave_add_u64(P.counters + CNT_VOXSTEPS, c.cVox);
        wave_add_u64(P.counters + CNT_HITS, c.cHits);
    }
}

__global__ __launch_bounds__(256) void k_collect_requested(const uint32_t* flags, uint32_t tiles, uint32_t* list,
                                                           uint32_t* count)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t < tiles && (flags[t] & 3u) == 3u) list[atomicAdd(count, 1u)] = t;
}

// Sets the visible bit of the listed tiles that are loaded.
__global__ __launch_bounds__(256) void k_mark_visible(uint32_t* flags, const uint32_t* list, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        const uint32_t t = list[i];
        uint32_t f = flags[t];
        if((f & 3u) == 2u && !(f & 4u)) flags[t] = f | 4u;
    }
}

// ---------------------------------------------------------------------------
// Batched lighting over many volumes in one persistent launch. Results do not
// depend on the work split: every voxel is lit by the same per-voxel code from
// the same inputs as in the single-volume pass.

// Per queue: units of its volumes (ceil(numRequests / unit)) and their
// exclusive prefix over the queue's volume list; resets the queue counters.
// One block of 1024.
__global__ __launch_bounds__(1024) void k_batch_prefix(DNbatchParams B)
{
    __shared__ uint32_t waveSums[16];
    __shared__ uint32_t carry;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    if(t < B.queues) { B.queueCounters[t] = 0u; B.directCounters[t] = 0u; }
    for(uint32_t q = 0; q < B.queues; q++)
    {
        const uint32_t len = B.queueLen[q];
        const uint32_t* ids = B.queueScenes + (size_t)q * B.numScenes;
        uint32_t* pre = B.queuePrefix + (size_t)q * (B.numScenes + 1u);
        if(t == 0) carry = 0u;
        __syncthreads();
        for(uint32_t base = 0; base < len; base += 1024u)
        {
            const uint32_t i = base + t;
            uint32_t u = 0;
            if(i < len) u = (*B.scenes[ids[i]].numRequests + B.unit - 1u) / B.unit;
            const uint32_t incl = wave_incl_scan(u);
            if(lane == 63u) waveSums[wave] = incl;
            __syncthreads();
            uint32_t before = carry;
            for(uint32_t w = 0; w < wave; w++) before += waveSums[w];
            if(i < len) pre[i] = before + incl - u;
            __syncthreads();
            if(t == 1023u) carry = before + incl;
            __syncthreads();
        }
        if(t == 0) pre[len] = carry;
        __syncthreads();
    }
}

__device__ __forceinline__ void flush_counters(uint64_t* cnt, RayCtx& c, uint32_t& lit)
{
    if(cnt)
    {
        wave_add_u64(cnt + CNT_RAYS, c.cRays);
        wave_add_u64(cnt + CNT_MAPSTEPS, c.cMap);
        wave_add_u64(cnt + CNT_VOXSTEPS, c.cVox);
        wave_add_u64(cnt + CNT_HITS, c.cHits);
        wave_add_u64(cnt + CNT_LIT, lit);
    }
    c.cRays = c.cMap = c.cVox = c.cHits = 0;
    lit = 0;
}

// Persistent: each wave takes units (`unit` requests of one volume, two
// requests per 64-lane pass) from its home queue, then from the others.
// Workgroups are dealt to the 8 XCDs round-robin by workgroup id, so with 8
// queues the home queue of workgroup b is b % 8 and a volume's units stay on
// one XCD (one L2) while its queue lasts.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(DN_LIGHT_WAVES, DN_LIGHT_WAVES)))
void k_lighting_batch(DNbatchParams B, const DNlightParams* __restrict__ scenes)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t home = blockIdx.x % B.queues;
    const bool dry = B.dry != 0u;
    __shared__ uint32_t maskRows[DN_MASK_LDS_WORDS(BLOCK)];
    RayCtx c;
    ctx_init(c);
    c.maskRow = lane_mask_row(maskRows);
    uint32_t lit = 0;
    uint32_t cur = 0xFFFFFFFFu;
    for(uint32_t k = 0; k < B.queues; k++)
    {
        const uint32_t q = (home + k) % B.queues;
        const uint32_t len = B.queueLen[q];
        if(len == 0) continue;
        const uint32_t* pre = B.queuePrefix + (size_t)q * (B.numScenes + 1u);
        const uint32_t total = pre[len];
        for(;;)
        {
            uint32_t u = 0;
            if(lane == 0) u = atomicAdd(&B.queueCounters[q], 1u);
            u = __builtin_amdgcn_readfirstlane(__shfl(u, 0, 64));
            if(u >= total) break;
            uint32_t lo = 0, hi = len - 1u;           // last volume with pre[i] <= u
            while(lo < hi)
            {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if(pre[mid] <= u) lo = mid;
                else hi = mid - 1u;
            }
            const uint32_t sid = __builtin_amdgcn_readfirstlane(B.queueScenes[(size_t)q * B.numScenes + lo]);
            if(sid != cur)
            {
                if(cur != 0xFFFFFFFFu && !dry) flush_counters(scenes[cur].counters, c, lit);
                cur = sid;
            }
            const DNlightParams& P = scenes[sid];
            DNsceneView sv = P.scene;                   // once per unit, into registers
            if(dry)
            {
                // compute only: no chunk requests, lastUsed stores into scratch
                sv.hasUnloaded = 0u;
                sv.lastUsed = B.dryLastUsed;
            }
            const uint32_t nreq = *P.numRequests;
            const uint32_t r0 = (u - pre[lo]) * B.unit;
            for(uint32_t j = 0; j < B.unit; j += 2u)
            {
                const uint32_t r = r0 + j + (lane >> 5);
                if(r < nreq) light_voxel(P, sv, r, lane & 31u, c, lit, dry);
            }
        }
    }
    if(cur != 0xFFFFFFFFu && !dry) flush_counters(scenes[cur].counters, c, lit);
}

// Copy-back and epilogue of every volume of a batch: blockIdx.y = volume.
__global__ __launch_bounds__(256) void k_batch_copy_back(const DNlightParams* scenes)
{
    const DNlightParams& P = scenes[blockIdx.y];
    if(!P.copyBack) return;
    copy_back_range(P.requests, *P.numRequests, P.scene.chunks, P.voxelsOut, (uint4*)P.scene.voxels,
                    blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u);
}

__global__ __launch_bounds__(256) void k_batch_epilogue(const DNlightParams* scenes)
{
    const DNlightParams& P = scenes[blockIdx.y];
    const uint32_t words = (P.tiles + 31u) / 32u;
    for(uint32_t w = blockIdx.x * 256u + threadIdx.x; w < words; w += gridDim.x * 256u)
        epilogue_word(P.scene.flags, (DNchunkRec*)P.scene.chunks, P.litBits, P.visProp, w);
}

int g_lightingBlocks = 0;
int g_cus = 0;

int cu_count()
{
    if(g_cus == 0)
    {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, dev);
        if(g_cus < 1) g_cus = 1;
    }
    return g_cus;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers

int dn_launch_draw(const DNdrawParams& p, hipStream_t stream)
{
    // persistent grid of resident blocks, no more than one wave per quad
    static int perCU = 0;
    if(perCU == 0)
    {
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_draw, 256, 0);
        if(perCU < 1) perCU = 1;
    }
    const uint32_t quads = ((p.width + 7u) / 8u) * ((p.height + 7u) / 8u);
    uint32_t blocks = (uint32_t)cu_count() * (uint32_t)perCU;
    if((quads + 3u) / 4u < blocks) blocks = (quads + 3u) / 4u;
    hipLaunchKernelGGL(k_draw, dim3(blocks), dim3(256), 0, stream, p);
    return (int)hipGetLastError();
}

int dn_launch_direct(const DNlightParams& p, uint32_t maxRequests, hipStream_t stream)
{
    if(p.numDirectRays == 0 || p.numDirectRays > 64u || !p.directOut) return (int)hipErrorInvalidValue;
    static int perCU = 0;
    if(perCU == 0)
    {
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_direct, 256, 0);
        if(perCU < 1) perCU = 1;
    }
    uint32_t blocks = (uint32_t)cu_count() * (uint32_t)perCU;
    if(maxRequests / 4 + 1 < blocks) blocks = maxRequests / 4 + 1;
    hipLaunchKernelGGL(k_direct, dim3(blocks), dim3(256), 0, stream, p);
    return (int)hipGetLastError();
}

int dn_launch_lighting(const DNlightParams& p, uint32_t maxRequests, hipStream_t stream)
{
    // persistent grid: as many blocks as can be resident with this map's
    // staged bitmap (recomputed when its size changes)
    static size_t lastOccBytes = ~(size_t)0;
    const size_t occBytes = sizeof(uint32_t) * ((p.tiles + 31u) / 32u);
    const bool lds = occBytes <= DN_OCC_LDS_MAX;
    if(g_lightingBlocks == 0 || occBytes != lastOccBytes)
    {
        int perCU = 0;
        if(lds) hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_lighting<true>, 256, occBytes);
        else hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_lighting<false>, 256, 0);
        if(perCU < 1) perCU = 1;
        g_lightingBlocks = cu_count() * perCU;
        lastOccBytes = occBytes;
    }
    // never more waves than two requests each
    uint32_t needWaves = (maxRequests + 1) / 2;
    uint32_t blocks = (uint32_t)g_lightingBlocks;
    if(needWaves / 4 + 1 < blocks) blocks = needWaves / 4 + 1;
    if(lds)
        hipLaunchKernelGGL(k_lighting<true>, dim3(blocks), dim3(256), occBytes, stream, p);
    else
        hipLaunchKernelGGL(k_lighting<false>, dim3(blocks), dim3(256), 0, stream, p);
    return (int)hipGetLastError();
}

int dn_launch_lighting_epilogue(const DNlightParams& p, uint32_t tiles, const uint32_t* requests,
                                const uint32_t* numRequests, int copyBack, hipStream_t stream)
{
    if(copyBack)
        hipLaunchKernelGGL(k_copy_back, dim3(1024), dim3(256), 0, stream, requests, numRequests, p.scene.chunks,
                           (const uint4*)p.voxelsOut, (uint4*)p.scene.voxels);
    uint32_t words = (tiles + 31) / 32;
    hipLaunchKernelGGL(k_light_epilogue, dim3((words + 255) / 256), dim3(256), 0, stream, p.scene.flags,
                       (DNchunkRec*)p.scene.chunks, p.litBits, p.visProp, words, p.workCounter);
    return (int)hipGetLastError();
}

int dn_launch_batch(const DNbatchParams& b, uint32_t blockThreads, uint32_t blocksPerCU, uint32_t maxTiles,
                    uint32_t maxRequests, bool epilogue, bool direct, hipStream_t stream)
{
    if(b.numScenes == 0) return 0;
    if(b.numScenes > 65535u || b.unit < 2u || (b.unit & 1u) || b.queues < 1u || b.queues > DN_BATCH_MAX_QUEUES ||
       (blockThreads != 64u && blockThreads != 256u))
        return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_prefix, dim3(1), dim3(1024), 0, stream, b);
    if(direct)
    {
        static int perCUd = 0;
        if(perCUd == 0)
        {
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCUd, k_direct_batch, 256, 0);
            if(perCUd < 1) perCUd = 1;
        }
        hipLaunchKernelGGL(k_direct_batch, dim3((uint32_t)cu_count() * (uint32_t)perCUd), dim3(256), 0, stream, b,
                           b.scenes);
    }
    static int perCU64 = 0, perCU256 = 0;
    int& perCU = blockThreads == 64u ? perCU64 : perCU256;
    if(perCU == 0)
    {
        if(blockThreads == 64u) hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_lighting_batch<64>, 64, 0);
        else hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_lighting_batch<256>, 256, 0);
        if(perCU < 1) perCU = 1;
    }
    uint32_t per = (uint32_t)perCU;
    if(blocksPerCU && blocksPerCU < per) per = blocksPerCU;
    const uint32_t blocks = (uint32_t)cu_count() * per;
    if(blockThreads == 64u) hipLaunchKernelGGL(k_lighting_batch<64>, dim3(blocks), dim3(64), 0, stream, b, b.scenes);
    else hipLaunchKernelGGL(k_lighting_batch<256>, dim3(blocks), dim3(256), 0, stream, b, b.scenes);
    if(epilogue && !b.dry)
    {
        uint32_t cx = (maxRequests * 32u + 255u) / 256u;
        cx = cx < 1u ? 1u : (cx > 64u ? 64u : cx);
        hipLaunchKernelGGL(k_batch_copy_back, dim3(cx, b.numScenes), dim3(256), 0, stream, b.scenes);
        uint32_t ex = ((maxTiles + 31u) / 32u + 255u) / 256u;
        ex = ex < 1u ? 1u : (ex > 64u ? 64u : ex);
        hipLaunchKernelGGL(k_batch_epilogue, dim3(ex, b.numScenes), dim3(256), 0, stream, b.scenes);
    }
    return (int)hipGetLastError();
}

int dn_launch_visible(const DNvisibleParams& p, hipStream_t stream)
{
    uint32_t blocks = (p.tiles + 1023) / 1024;
    if(blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_visible_count, dim3(blocks), dim3(1024), 0, stream, p);
    hipLaunchKernelGGL(k_visible_write, dim3(blocks), dim3(1024), 0, stream, p);
    return (int)hipGetLastError();
}

int dn_launch_upload(uint32_t* flags, uint32_t* lastUsed, DNchunkRec* chunks, uint4* voxels, uint32_t* fresh,
                     const DNuploadOp* ops, uint32_t numOps, const DNchunkRec* recs, const uint4* stagedVoxels,
                     hipStream_t stream)
{
    if(numOps == 0) return 0;
    hipLaunchKernelGGL(k_upload, dim3(numOps), dim3(256), 0, stream, flags, lastUsed, chunks, voxels, fresh, ops, recs,
                       stagedVoxels);
    return (int)hipGetLastError();
}

int dn_launch_rays(const DNrayParams& p, hipStream_t stream)
{
    if(p.n == 0) return 0;
    hipLaunchKernelGGL(k_rays, dim3((p.n + 255) / 256), dim3(256), 0, stream, p);
    return (int)hipGetLastError();
}

int dn_launch_collect_requested(const uint32_t* flags, uint32_t tiles, uint32_t* list, uint32_t* count,
                                hipStream_t stream)
{
    if(tiles == 0) return 0;
    hipLaunchKernelGGL(k_collect_requested, dim3((tiles + 255) / 256), dim3(256), 0, stream, flags, tiles, list, count);
    return (int)hipGetLastError();
}

int dn_launch_mark_visible(uint32_t* flags, const uint32_t* list, uint32_t n, hipStream_t stream)
{
    if(n == 0) return 0;
    hipLaunchKernelGGL(k_mark_visible, dim3((n + 255) / 256), dim3(256), 0, stream, flags, list, n);
    return (int)hipGetLastError();
}

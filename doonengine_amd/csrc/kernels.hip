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

__global__ __launch_bounds__(256) void k_draw(DNdrawParams P)
{
    const DNsceneView& s = P.scene;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t x = blockIdx.x * 16u + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t y = blockIdx.y * 16u + (wave >> 1) * 8u + (lane >> 3);
    RayCtx c;
    ctx_init(c);
    uint32_t drawn = 0;
    if(x < P.width && y < P.height)
    {
        drawn = 1;
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
            bool hit = step_map<true>(s, c, dir, inv, p, false, nrm, vox, ca, cm);
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
    if(P.counters)
    {
        wave_add_u64(P.counters + CNT_RAYS, c.cRays);
        wave_add_u64(P.counters + CNT_MAPSTEPS, c.cMap);
        wave_add_u64(P.counters + CNT_VOXSTEPS, c.cVox);
        wave_add_u64(P.counters + CNT_HITS, c.cHits);
        wave_add_u64(P.counters + CNT_PIXELS, drawn);
    }
}

// ---------------------------------------------------------------------------
// lighting pass (assets/shaders/voxelLighting.comp:65-288). One lane per stored
// voxel; a wave takes two consecutive 32-voxel requests (normally the two
// halves of one chunk's voxel range) from a device work counter.

// One lighting ray: step_map(dir, inv = 1/dir, pos, ignoreFirst = true, ...) as
// voxelLighting.comp calls it. The nested chunk-in-map loops are measured
// faster here than the single flattened loop (DN_LIGHT_FLAT); both give the
// same results bit for bit.
__device__ __forceinline__ bool light_trace(const DNsceneView& s, RayCtx& c, f3 dir, f3 inv, f3& p, DVoxel& vox,
                                            f3& ca, float& cm)
{
    f3 tn = splat(0.0f);
#ifdef DN_LIGHT_FLAT
    return step_map_flat<false>(s, c, dir, inv, p, true, vox, ca, cm, tn);
#else
    return step_map<false>(s, c, dir, inv, p, true, tn, vox, ca, cm);
#endif
}

struct LCtx {
    const DNlightParams* P;
    RayCtx* c;
    bool firstSample;
    f3 sunDir, sunStrength;
};

__device__ void shadow_ray(LCtx& L, f3 p, uint32_t sample, f3& color)
{
    const DNlightParams& P = *L.P;
    DVoxel hv = {};
    f3 ca;
    float cm;
    if(P.numDirectRays == 0)
    {
        const float* r = L.firstSample ? P.firstShadow : P.shadowRays[sample];
        f3 pp = p;
        if(!light_trace(P.scene, *L.c, ld3(r), ld3(r + 3), pp, hv, ca, cm))
            color = color + (L.sunStrength * cm + ca);
        return;
    }
    f3 acc = splat(0.0f);
    for(uint32_t k = 0; k < P.numDirectRays; k++)
    {
        f3 pp = p;
        if(!light_trace(P.scene, *L.c, ld3(P.directRays[k]), ld3(P.directRays[k] + 3), pp, hv, ca, cm))
            acc = acc + (L.sunStrength * cm + ca);
    }
    color = color + acc / (float)P.numDirectRays;
}

__device__ void specular_ray(LCtx& L, uint32_t thisMapIndex, f3 p, f3 dir, f3 albedo, uint32_t reflectType, f3& color)
{
    const DNlightParams& P = *L.P;
    const DNsceneView& s = P.scene;
    DVoxel hv = {};
    f3 lastPos = p;
    f3 mult = albedo;
    for(uint32_t i = 0; i < P.specBounceLimit; i++)
    {
        f3 ca; float cm;
        if(light_trace(s, *L.c, dir, vinv(dir), p, hv, ca, cm))
        {
            int32_t hx = (int32_t)p.x, hy = (int32_t)p.y, hz = (int32_t)p.z;
            uint32_t hitIdx = (uint32_t)hx + s.mapX * ((uint32_t)hy + s.mapY * (uint32_t)hz);
            bool inMap = hx >= 0 && hy >= 0 && hz >= 0 && (uint32_t)hx < s.mapX && (uint32_t)hy < s.mapY &&
                         (uint32_t)hz < s.mapZ;
            if(inMap && (s.flags[thisMapIndex] & 4u))
            {
                uint32_t bit = 1u << (hitIdx & 31u);
                if(!(P.visProp[hitIdx >> 5] & bit)) atomicOr(&P.visProp[hitIdx >> 5], bit);
            }
            f3 dist = vabs(vfloor(p * 8.0f) - vfloor(lastPos * 8.0f));
            if(dot3(dist, dist) <= 1.0f) return;
            const DNmat& hm = s.mats[hv.material];
            hv.diffuse = hv.diffuse * (1.0f - hm.specular);
            if(hm.emissive)
            {
                color = color + ((hv.albedo * emission_of(hm)) * cm + ca) * mult * albedo;
                return;
            }
            f3 hitColor = hv.diffuse * hv.albedo;
            color = color + (hitColor * cm + ca) * mult;
            if(hm.specular == 0.0f) return;
            mult = mult * ((hv.albedo * cm) * hm.specular);
            reflectType = hm.reflectType;
            lastPos = p;
            dir = vreflect(dir, hv.normal);
        }
        else if(dot3(dir, L.sunDir) > 0.99f)
        {
            color = color + (L.sunStrength * cm + ca);
            return;
        }
        else
        {
            f3 src = reflectType == 1u ? sky_color(s, dir) : L.sunStrength;
            color = color + (src * cm + ca) * mult;
            return;
        }
    }
}

__device__ void diffuse_ray(LCtx& L, f3 normal, f3 p, const DVoxel& initial, uint32_t sample, f3& color)
{
    const DNlightParams& P = *L.P;
    const DNsceneView& s = P.scene;
    DVoxel hv = initial;
    f3 hitNormal = normal;
    float hitSpec = 0.0f;
    uint32_t hitShine = 0;
    f3 newColor = splat(1.0f);
    f3 lastPos = p;
    f3 lastDir = splat(0.0f);
    for(uint32_t i = 0; i < P.diffuseBounceLimit; i++)
    {
        f3 dir;
        if(i > 0 && (P.rnd.glossyPick[sample][i] + 1.0f) * 0.5f < hitSpec)
            dir = vnorm(vreflect(lastDir, hitNormal) * (float)hitShine + ld3(P.rnd.glossyDir[i]));
        else if(L.firstSample)
            dir = vnorm(hitNormal) + DN_EPS;
        else
            dir = vnorm(hitNormal + ld3(P.rnd.diffuseDir[sample][i])) + DN_EPS;
        f3 ca;
        float cm;
        DVoxel nv = {};
        bool hit = light_trace(s, *L.c, dir, vinv(dir), p, nv, ca, cm);
        if(hit)
        {
            hv = nv;
            hitNormal = hv.normal;
            const DNmat& hm = s.mats[hv.material];
            hitSpec = hm.specular;
            hitShine = hm.shininess;
            f3 dist = vabs(vfloor(lastPos * 8.0f) - vfloor(p * 8.0f));
            if(dot3(dist, dist) < 1.0f) return;
            if(hm.emissive)
            {
                color = color + newColor * ((hv.albedo * emission_of(hm)) * cm + ca);
                return;
            }
            newColor = newColor * (hv.albedo * cm + ca);
        }
        else
        {
            float cosv = gmax(dot3(dir, L.sunDir), 0.0f);
            color = color + (((newColor * cosv) * L.sunStrength) * cm + ca);
            return;
        }
        lastDir = dir;
    }
}

__device__ __forceinline__ uint32_t rne_u(float x) { return (uint32_t)rintf(x); }

__device__ void light_voxel(const DNlightParams& P, uint32_t r, uint32_t lane32, RayCtx& c, uint32_t& lit)
{
    const DNsceneView& s = P.scene;
    const uint32_t req = P.requests[r];
    const uint32_t mapIndex = req >> 4, block = req & 15u;
    const uint32_t voxNum = lane32 + block * 32u;
    const DNchunkRec* rec = s.chunks + mapIndex;
    const size_t slot = (size_t)r * 32u + lane32;
    int32_t sel = select_in(rec->mask, voxNum);
    const bool dbg = slot < P.debugSlots;
    if(sel < 0)
    {
        if(dbg && P.debugIndex) P.debugIndex[slot] = 0xFFFFFFFFu;
        return;
    }
    if(voxNum == 0)
        atomicOr(&P.litBits[mapIndex >> 5], 1u << (mapIndex & 31u));
    c.lastVoxID = 255u; c.lastVoxRefract = 1.0f;
    const int32_t cx = sel % 8, cy = (sel / 8) % 8, cz = sel / 64;
    const uint32_t vi = rec->voxelIndex + voxNum;
    const uint4 cw = s.voxels[vi];
    const DVoxel me = decompress(cw);
    const DNmat& mat = s.mats[me.material];
    const uint32_t ns = rec->samples;
    const float indirect = (float)(ns < P.maxDiffuseSamples ? ns : P.maxDiffuseSamples);
    LCtx L;
    L.P = &P; L.c = &c;
    L.firstSample = indirect == 0.0f;
    L.sunDir = ld3(P.sunDir);
    L.sunStrength = ld3(s.sunStrength);
    f3 p = (mk((float)cx, (float)cy, (float)cz) * 0.125f + mk((float)rec->pos[0], (float)rec->pos[1], (float)rec->pos[2])) +
           splat(0.0625f);
    p = p + splat(0.0625f - DN_EPS) * me.normal;
    f3 spec = splat(0.0f), diff = splat(0.0f);
    f3 view = p - ld3(P.camPos);
    if(mat.specular > 0.0f && dot3(view, me.normal) < 0.0f && mat.reflectType <= 1u)
    {
        f3 refl = vreflect(vnorm(view), me.normal);
        for(uint32_t i = 0; i < P.numSpecRays; i++)
        {
            f3 sd = vnorm(refl * (float)mat.shininess + ld3(P.specPoints + 3 * i)) + DN_EPS;
            specular_ray(L, mapIndex, p, sd, me.albedo, mat.reflectType, spec);
        }
        spec = spec / (float)P.numSpecRays;
    }
    if(mat.specular < 1.0f)
    {
        for(uint32_t i = 0; i < P.numDiffuseSamples; i++)
        {
            diff = diff + ld3(P.ambient);
            diffuse_ray(L, me.normal + DN_EPS, p, me, i, diff);
            shadow_ray(L, p, i, diff);
        }
        diff = (me.diffuse * indirect + diff) / (indirect + (float)P.numDiffuseSamples);
    }
    if(dbg && P.debugValues)
    {
        float* dv = P.debugValues + slot * 6;
        dv[0] = diff.x; dv[1] = diff.y; dv[2] = diff.z; dv[3] = spec.x; dv[4] = spec.y; dv[5] = spec.z;
    }
    if(dbg && P.debugIndex) P.debugIndex[slot] = vi;
    spec = vclamp(spec, 0.0f, 1.0f);
    diff = vclamp(diff, 0.0f, 1.0f);
    const uint32_t wdx = rne_u(diff.x * 65535.0f), wdy = rne_u(diff.y * 65535.0f), wdz = rne_u(diff.z * 65535.0f);
    const uint32_t ar = rne_u(me.albedo.x * 255.0f) & 255u, ag = rne_u(me.albedo.y * 255.0f) & 255u,
                   ab = rne_u(me.albedo.z * 255.0f) & 255u;
    const uint32_t sx = rne_u(spec.x * 255.0f) & 255u, sy = rne_u(spec.y * 255.0f) & 255u,
                   sz = rne_u(spec.z * 255.0f) & 255u;
    uint4 o;
    o.x = cw.x;
    o.y = (ar << 24) | (ag << 16) | (ab << 8) | sx;
    o.z = (sy << 24) | (sz << 16) | (wdx & 0xFFFFu);
    o.w = ((wdy & 0xFFFFu) << 16) | (wdz & 0xFFFFu);
    P.voxelsOut[vi] = o;
    lit++;
}

__global__ __launch_bounds__(256) void k_lighting(DNlightParams P)
{
    const uint32_t lane = threadIdx.x & 63u;
    RayCtx c;
    ctx_init(c);
    uint32_t lit = 0;
    const uint32_t nreq = *P.numRequests;
    for(;;)
    {
        uint32_t base = 0;
        if(lane == 0) base = atomicAdd(P.workCounter, 2u);
        base = __shfl(base, 0, 64);
        if(base >= nreq) break;
        const uint32_t r = base + (lane >> 5);
        if(r < nreq) light_voxel(P, r, lane & 31u, c, lit);
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

// Pass epilogue over the tile bitmaps: clear the visible bit and bump the
// sample count of lit chunks, then set the visible bit of chunks reached by
// specular rays from visible chunks; reset the work counter.
__global__ __launch_bounds__(256) void k_light_epilogue(uint32_t* flags, DNchunkRec* chunks, uint32_t* litBits,
                                                        uint32_t* visProp, uint32_t words, uint32_t* workCounter)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if(w == 0) *workCounter = 0u;
    if(w >= words) return;
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

// Copies the lit voxels of this pass from the output pool back into the pool
// (only used when the pass wrote to a separate buffer).
__global__ __launch_bounds__(256) void k_copy_back(const uint32_t* requests, const uint32_t* numRequests,
                                                   const DNchunkRec* chunks, const uint4* src, uint4* dst)
{
    const uint32_t n = *numRequests;
    for(uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n * 32u; g += gridDim.x * blockDim.x)
    {
        uint32_t req = requests[g >> 5];
        uint32_t voxNum = (g & 31u) + (req & 15u) * 32u;
        const DNchunkRec* rec = chunks + (req >> 4);
        if(voxNum < rec->numVoxels) dst[rec->voxelIndex + voxNum] = src[rec->voxelIndex + voxNum];
    }
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

int g_lightingBlocks = 0;

}  // namespace

// ---------------------------------------------------------------------------
// launchers

int dn_launch_draw(const DNdrawParams& p, hipStream_t stream)
{
    dim3 grid((p.width + 15) / 16, (p.height + 15) / 16);
    hipLaunchKernelGGL(k_draw, grid, dim3(256), 0, stream, p);
    return (int)hipGetLastError();
}

int dn_launch_lighting(const DNlightParams& p, uint32_t maxRequests, hipStream_t stream)
{
    if(g_lightingBlocks == 0)
    {
        int dev = 0, cus = 0, perCU = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_lighting, 256, 0);
        if(perCU < 1) perCU = 1;
        g_lightingBlocks = cus * perCU;
        if(g_lightingBlocks < 1) g_lightingBlocks = 1;
    }
    // never more waves than two requests each
    uint32_t needWaves = (maxRequests + 1) / 2;
    uint32_t blocks = (uint32_t)g_lightingBlocks;
    if(needWaves / 4 + 1 < blocks) blocks = needWaves / 4 + 1;
    hipLaunchKernelGGL(k_lighting, dim3(blocks), dim3(256), 0, stream, p);
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

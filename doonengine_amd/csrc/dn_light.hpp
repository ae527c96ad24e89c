// Device functions of the lighting pass (assets/shaders/voxelLighting.comp:65-288),
// shared by the single-volume, batched and radiance-cache kernels.
#pragma once
#include "dn_device.hpp"

namespace {

// One lighting ray: step_map(dir, inv = 1/dir, pos, ignoreFirst = true, ...) as
// voxelLighting.comp calls it, with the chunk masks in the lane's LDS row
// (c.maskRow, set by every kernel that traces lighting rays). The nested
// chunk-in-map loops are measured faster here than the single flattened loop.
__device__ __forceinline__ bool light_trace(const DNsceneView& s, RayCtx& c, f3 dir, f3 inv, f3& p, DVoxel& vox,
                                            f3& ca, float& cm)
{
    f3 tn = splat(0.0f);
    return step_map<false, true>(s, c, dir, inv, p, true, tn, vox, ca, cm);
}

// light_trace for the frame-uniform shadow directions (the same for every
// voxel, so the test is uniform across the wave): the finite-delta walk when
// no direction component is zero.
__device__ __forceinline__ bool light_trace_uniform(const DNsceneView& s, RayCtx& c, f3 dir, f3 inv, f3& p,
                                                    DVoxel& vox, f3& ca, float& cm)
{
    f3 tn = splat(0.0f);
    if(all_finite(inv)) return step_map<false, true, true>(s, c, dir, inv, p, true, tn, vox, ca, cm);
    return step_map<false, true>(s, c, dir, inv, p, true, tn, vox, ca, cm);
}

// The calling lane's row of a block's mask rows.
__device__ __forceinline__ uint32_t* lane_mask_row(uint32_t* rows)
{
    return rows + threadIdx.x * DN_MASK_ROW;
}

// Per-voxel context. `s` is the scene view every ray walks: the caller keeps it
// in registers (a kernel argument, or a copy made once per work unit when the
// parameters live in memory), so the DDA loops never reload it.
struct LCtx {
    const DNlightParams* P;
    const DNsceneView* s;
    RayCtx* c;
    bool firstSample;
    bool dry;                    // compute only: no stores, no bitmap updates
    f3 sunDir, sunStrength;
    const DNdirectTerm* direct;  // this voxel's direct-ray term traced ahead, or null
};

// Start point of the lighting rays of stored voxel `sel` of a chunk
// (voxelLighting.comp:231-232): the voxel centre moved 1/16 - EPS along its normal.
__device__ __forceinline__ f3 voxel_origin(const DNchunkRec* rec, int32_t sel, f3 normal)
{
    const int32_t cx = sel % 8, cy = (sel / 8) % 8, cz = sel / 64;
    f3 p = (mk((float)cx, (float)cy, (float)cz) * 0.125f + mk((float)rec->pos[0], (float)rec->pos[1], (float)rec->pos[2])) +
           splat(0.0625f);
    return p + splat(0.0625f - DN_EPS) * normal;
}

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
        if(!light_trace(*L.s, *L.c, ld3(r), ld3(r + 3), pp, hv, ca, cm))
            color = color + (L.sunStrength * cm + ca);
        return;
    }
    // the terms traced ahead assume every ray starts with lastVoxID = 255 and
    // ends so (w = 0 when they all did)
    if(L.direct && L.c->lastVoxID == 255u)
    {
        const float4 d = L.direct->term;
        if(d.w == 0.0f)
        {
            const uint4 n = L.direct->counts;
            L.c->cRays += n.x; L.c->cMap += n.y; L.c->cVox += n.z; L.c->cHits += n.w;
            color = color + mk(d.x, d.y, d.z);
            return;
        }
    }
    f3 acc = splat(0.0f);
    for(uint32_t k = 0; k < P.numDirectRays; k++)
    {
        f3 pp = p;
        if(!light_trace_uniform(*L.s, *L.c, ld3(P.directRays[k]), ld3(P.directRays[k] + 3), pp, hv, ca, cm))
            acc = acc + (L.sunStrength * cm + ca);
    }
    color = color + acc / (float)P.numDirectRays;
}

__device__ void specular_ray(LCtx& L, uint32_t thisMapIndex, f3 p, f3 dir, f3 albedo, uint32_t reflectType, f3& color)
{
    const DNlightParams& P = *L.P;
    const DNsceneView& s = *L.s;
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
            if(inMap && !L.dry && (s.flags[thisMapIndex] & 4u))
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

// Random numbers of the diffuse path as the shader draws them: the same for
// every voxel of a frame (seeds depend on time, sample and bounce only).
struct FrameRandom {
    const DNframeRandom* r;
    __device__ float glossy_pick(uint32_t sample, uint32_t i) const { return r->glossyPick[sample][i]; }
    __device__ f3 glossy_dir(uint32_t, uint32_t i) const { return ld3(r->glossyDir[i]); }
    __device__ f3 diffuse_dir(uint32_t sample, uint32_t i) const { return ld3(r->diffuseDir[sample][i]); }
};

template <typename RNG>
__device__ void diffuse_ray(LCtx& L, f3 normal, f3 p, const DVoxel& initial, uint32_t sample, f3& color,
                            const RNG& rng)
{
    const DNlightParams& P = *L.P;
    const DNsceneView& s = *L.s;
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
        if(i > 0 && (rng.glossy_pick(sample, i) + 1.0f) * 0.5f < hitSpec)
            dir = vnorm(vreflect(lastDir, hitNormal) * (float)hitShine + rng.glossy_dir(sample, i));
        else if(L.firstSample)
            dir = vnorm(hitNormal) + DN_EPS;
        else
            dir = vnorm(hitNormal + rng.diffuse_dir(sample, i)) + DN_EPS;
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

__device__ void light_voxel(const DNlightParams& P, const DNsceneView& s, uint32_t r, uint32_t lane32, RayCtx& c,
                            uint32_t& lit, bool dry)
{
    const uint32_t req = P.requests[r];
    const uint32_t mapIndex = req >> 4, block = req & 15u;
    const uint32_t voxNum = lane32 + block * 32u;
    const DNchunkRec* rec = s.chunks + mapIndex;
    const size_t slot = (size_t)r * 32u + lane32;
    int32_t sel = select_in(rec->mask, voxNum);
    const bool dbg = slot < P.debugSlots && !dry;
    if(sel < 0)
    {
        if(dbg && P.debugIndex) P.debugIndex[slot] = 0xFFFFFFFFu;
        if(!dry && slot < P.freshSlots && P.debugFresh)
            reinterpret_cast<float4*>(P.debugFresh)[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    if(voxNum == 0 && !dry)
        atomicOr(&P.litBits[mapIndex >> 5], 1u << (mapIndex & 31u));
    c.lastVoxID = 255u; c.lastVoxRefract = 1.0f;
    const uint32_t vi = rec->voxelIndex + voxNum;
    const uint4 cw = s.voxels[vi];
    const DVoxel me = decompress(cw);
    const DNmat& mat = s.mats[me.material];
    const uint32_t ns = rec->samples;
    const float indirect = (float)(ns < P.maxDiffuseSamples ? ns : P.maxDiffuseSamples);
    LCtx L;
    L.P = &P; L.s = &s; L.c = &c; L.dry = dry;
    L.firstSample = indirect == 0.0f;
    L.sunDir = ld3(P.sunDir);
    L.sunStrength = ld3(s.sunStrength);
    L.direct = (P.directIn && P.numDirectRays && slot < P.directSlots) ? P.directIn + slot : nullptr;
    f3 p = voxel_origin(rec, sel, me.normal);
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
    f3 fresh = splat(0.0f);
    if(mat.specular < 1.0f)
    {
        const FrameRandom rng{&P.rnd};
        for(uint32_t i = 0; i < P.numDiffuseSamples; i++)
        {
            diff = diff + ld3(P.ambient);
            diffuse_ray(L, me.normal + DN_EPS, p, me, i, diff, rng);
            shadow_ray(L, p, i, diff);
        }
        if(P.numDiffuseSamples) fresh = diff / (float)P.numDiffuseSamples;
        diff = (me.diffuse * indirect + diff) / (indirect + (float)P.numDiffuseSamples);
    }
    if(dbg && P.debugValues)
    {
        float* dv = P.debugValues + slot * 6;
        dv[0] = diff.x; dv[1] = diff.y; dv[2] = diff.z; dv[3] = spec.x; dv[4] = spec.y; dv[5] = spec.z;
    }
    if(!dry && slot < P.freshSlots && P.debugFresh)
        reinterpret_cast<float4*>(P.debugFresh)[slot] = make_float4(fresh.x, fresh.y, fresh.z, 1.0f);
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
    if(!dry) P.voxelsOut[vi] = o;
    lit++;
}


}  // namespace

// Internal layouts and kernel interfaces shared by the host engine and the HIP
// kernels. Nothing here is part of the public C-ABI.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <hip/hip_runtime.h>

// One record per map tile (indexed like the reference's chunk buffer, by map
// index). 128 bytes = one cache line: the occupancy mask, the rank prefix and
// the pool offset that a DDA needs on chunk entry and on a hit arrive together.
struct alignas(128) DNchunkRec {
    uint32_t mask[16];     // bit (x + 8y + 64z): voxel stored (voxelShared.comp:45)
    uint32_t partial[3];   // stored voxels before bits 128, 256, 384 (voxelShared.comp:44)
    uint32_t voxelIndex;   // pool start of this chunk's voxels (voxelShared.comp:53)
    uint32_t numVoxels;    // stored voxels (popcount of mask)
    uint32_t samples;      // numIndirectSamples (voxelShared.comp:42)
    int32_t pos[3];        // chunk position in the map (voxelShared.comp:41)
    uint32_t pad[7];
};
static_assert(sizeof(DNchunkRec) == 128, "chunk record must be one 128-byte line");

// Material, std140 32 bytes (voxelShared.comp:57-69, voxel.h:82-94).
struct DNmat {
    float padding[2];      // padding[0] > 0: emission strength (extension)
    uint32_t emissive;
    float opacity;
    float refractIndex;
    float specular;
    uint32_t reflectType;
    uint32_t shininess;
};
static_assert(sizeof(DNmat) == 32, "material layout");

// Read-only view of one volume's device state used by every DDA.
struct DNsceneView {
    const uint32_t* occLoaded;   // bit per tile: flags & 3 == 2
    const uint32_t* occPresent;  // bit per tile: flags & 3 in {1, 3}
    uint32_t* flags;             // per tile: load state (bits 0-1), visible (bit 2)
    uint32_t* lastUsed;          // per tile
    const DNchunkRec* chunks;    // per tile
    const uint4* voxels;         // pool
    const DNmat* mats;           // 256
    uint32_t mapX, mapY, mapZ;
    uint32_t hasUnloaded;        // any tile in state 1/3: rays may request chunks
    float sunStrength[3];
    float skyBot[3], skyTop[3];
    const uchar4* cube;          // sky cubemap: 6 faces (+X, -X, +Y, -Y, +Z, -Z) of cubeSize^2 RGBA8, or null
    uint32_t cubeSize;
    // Box of the present tiles (flags & 3 != 0, loaded or not), inclusive, in
    // chunk units; lo > hi when there is none. Lighting rays stop once beyond
    // it on the side they move toward: no voxel, request or stamp lies there.
    int32_t presLo[3], presHi[3];
    // Four tables of mapZ x mapX, [q][z][x]: the highest present tile's y over
    // the columns a ray in column (x, z) can still reach when it moves in
    // quadrant q (bit 0: x steps up, else x stays or steps down; bit 1: the
    // same for z), -1 where there is none. A lighting ray that does not move
    // down stops once above that height (or null: no such rule).
    const int32_t* colTop;
};

// Limits of the per-frame tables passed by value to the lighting kernel.
enum { DN_MAX_POINTS = 64, DN_MAX_SAMPLES = 4, DN_MAX_BOUNCES = 8 };

// Per-frame voxel-uniform random numbers (voxelLighting.comp:29-57): every seed
// in the shader depends only on time, the sample index and the bounce index,
// so all voxels of a frame draw the same numbers. They are computed once on
// the host with the shader's hash and passed to the kernel.
struct DNframeRandom {
    float diffuseDir[DN_MAX_SAMPLES][DN_MAX_BOUNCES][3];  // rand_unit_sphere(time*(i+1) + j)
    float glossyPick[DN_MAX_SAMPLES][DN_MAX_BOUNCES];     // rand(time*(i+1) + bounceLimit + j)
    float glossyDir[DN_MAX_BOUNCES][3];                   // rand_unit_sphere(time + j)
    float shadowDir[DN_MAX_SAMPLES][3];                   // rand_unit_sphere(time*(i+1+numSamples))
};

// What a voxel's direct rays add per diffuse sample (term.xyz), whether that
// value holds (term.w == 0: every ray started and ended with lastVoxID = 255),
// and the ray counters the rays produced (rays, chunk-grid trips, voxel trips,
// hits), which the pass adds each time it uses the value.
struct DNdirectTerm {
    float4 term;
    uint4 counts;
};

struct DNlightParams {
    DNsceneView scene;
    const uint32_t* requests;
    const uint32_t* numRequests;   // device count
    uint32_t* workCounter;         // device, reset by the epilogue
    uint32_t* litBits;             // bit per tile, set for lit chunks
    uint32_t* visProp;             // bit per tile, set by specular rays
    uint4* voxelsOut;              // pool receiving lit words (== voxels when written in place)
    float* debugValues;            // optional: 6 floats per slot
    uint32_t* debugIndex;          // optional: pool index per slot
    float* debugFresh;             // optional: 4 floats per slot, this pass's diffuse samples
                                   // averaged before the running mean (w = 1 where written)
    uint32_t debugSlots;           // capacity of the debug outputs in slots
    uint32_t freshSlots;           // capacity of debugFresh in slots
    uint64_t* counters;            // 8 x u64
    uint32_t numDiffuseSamples, maxDiffuseSamples, diffuseBounceLimit, specBounceLimit;
    float sunDir[3];
    float ambient[3];
    float shadowSoftness;
    float camPos[3];
    uint32_t numSpecRays, numDirectRays;
    const float* specPoints;       // device, numSpecRays * 3
    DNframeRandom rnd;
    // Shadow-ray directions are the same for every voxel of a frame, so they
    // come precomputed with their inverses (dir xyz, 1/dir xyz), in the exact
    // FP32 operations of voxelLighting.comp:72-75:
    float directRays[DN_MAX_POINTS][6];   // normalize(sunDir * softness + P_k) + eps, k < numDirectRays
    float shadowRays[DN_MAX_SAMPLES][6];  // normalize(sunDir * softness + rand_unit_sphere(seed_i)) + eps
    float firstShadow[6];                 // sunDir + eps (first sample)
    uint32_t tiles;                       // map tiles (bitmap words = ceil(tiles / 32))
    uint32_t copyBack;                    // voxelsOut != voxels: copy lit voxels back after the pass
    // Direct-ray terms traced ahead of the pass (k_direct), one per request
    // slot. Null: the pass traces them itself.
    DNdirectTerm* directOut;              // written by k_direct
    const DNdirectTerm* directIn;         // read by the pass
    uint32_t directSlots;                 // capacity of both in request slots
};

// Batched lighting over many volumes in one persistent launch. Work units are
// `unit` consecutive requests of one volume; volumes are dealt to `queues`
// queues (1, or 8 = one per XCD), each with its own unit counter; a wave drains
// its own XCD's queue first, then the others.
enum { DN_BATCH_MAX_QUEUES = 8 };
struct DNbatchParams {
    const DNlightParams* scenes;   // device, one per volume
    uint32_t numScenes;
    uint32_t unit;                 // requests per work unit, even
    uint32_t queues;               // 1 or 8
    uint32_t dry;                  // 1: compute only, no stores (autotuning)
    uint32_t* dryLastUsed;         // dry: where rays' lastUsed stores go instead (>= the largest volume's tiles)
    uint32_t* queueCounters;       // [queues] next unit of each queue
    uint32_t* directCounters;      // [queues] next unit of each queue for the direct-ray pass
    const uint32_t* queueScenes;   // [queues][numScenes] volume ids of each queue, in order
    const uint32_t* queueLen;      // [queues] volumes in each queue
    uint32_t* queuePrefix;         // [queues][numScenes + 1] exclusive unit prefix, written on device
};

struct DNdrawParams {
    DNsceneView scene;
    float invView[16];
    float invViewProj[16];
    uint32_t width, height;
    uint32_t viewMode;
    float4* out;
    void* hitOut;                  // optional DNXpixelHit*
    uint64_t* counters;
    uint32_t* quadCounter;         // device [2]: next 8x8 quad, finished waves; zero between launches
};

struct DNrayParams {
    DNsceneView scene;
    uint32_t n;
    const float* origins;
    const float* dirs;
    uint32_t ignoreFirst;
    void* out;                     // DNXrayHit*
    uint64_t* counters;
};

// One upload operation of a sync: a tile state change, optionally with a chunk
// record and its voxels staged in device memory.
struct DNuploadOp {
    uint32_t tile;
    uint32_t mode;        // 0: flags = 0, 1: flags = 1, 2: load record + voxels, flags = 2 | visible
    uint32_t recSlot;     // index into staged records (mode 2)
    uint32_t voxOffset;   // first staged voxel (mode 2)
};

struct DNvisibleParams {
    uint32_t* flags;
    uint32_t* lastUsed;
    const DNchunkRec* chunks;
    uint32_t* fresh;       // bit per tile
    uint32_t tiles;
    uint32_t split;        // 0: emit nothing (DN_WRITE)
    uint32_t frameNum;
    uint32_t* blockSums;
    uint32_t* requests;
    uint32_t* numRequests;
};

// Counter slots (u64 each) per kind.
enum { CNT_RAYS = 0, CNT_MAPSTEPS, CNT_VOXSTEPS, CNT_HITS, CNT_LIT, CNT_PIXELS, CNT_SLOTS = 8 };

// Device sphere edits (edits.hip). A batch of spheres is split into (tile,
// sphere) pairs, numbered in the reference's order (sphere by sphere, tiles of
// a sphere's box in z, y, x order); each touched tile lists its pairs in that
// order. Per pair the kernel reports what the edit did to the tile's chunk.
enum { DN_EDIT_MODIFIED = 1, DN_EDIT_ADDED = 2, DN_EDIT_REMOVED = 4 };
struct DNeditSphere {
    float c[3];
    float r2, r12;                 // r*r and (r+1)*(r+1), FP32
    uint32_t material;             // 255: carve
    uint32_t flip;                 // carve: turn the next shell's normals inwards
    uint32_t carveNormal;          // carve: the caller's voxel, compressed
    uint32_t albedo;               // the caller's albedo word (fill and carve)
};
struct DNeditParams {
    const DNeditSphere* spheres;
    const uint32_t* tileIds;       // [tiles of the batch] map index
    const uint32_t* tileInfo;      // [tiles of the batch] chunk present (bit 31) | stored voxels
    const uint32_t* pairStart;     // [tiles of the batch + 1] into pairSphere / pairSeq
    const uint32_t* pairSphere;    // sphere of each pair, in edit order per tile
    const uint32_t* pairSeq;       // sequence number of each pair
    uint8_t* events;               // [pairs] DN_EDIT_* by sequence number
    uint32_t* tileOut;             // [tiles of the batch] present (bit 31) | modified (bit 30) | stored voxels
    uint2* words;                  // device copy of the CPU chunk words, [map tiles][512], voxel x*64 + y*8 + z
    uint32_t mapX, mapY, mapZ;
};
int dn_launch_sphere_edits(const DNeditParams& p, uint32_t numTiles, hipStream_t stream);
int dn_launch_edit_scatter(uint2* words, const uint32_t* tiles, const uint2* in, uint32_t n, hipStream_t stream);
int dn_launch_edit_gather(const uint2* words, const uint32_t* tiles, uint2* out, uint32_t n, hipStream_t stream);

// Fills the lighting-pass parameters of `vol` for a frame (engine.cpp).
struct DNvolume;
bool prepare_lighting(DNvolume* vol, int numDiffuseSamples, int maxDiffuseSamples, float time, DNlightParams& p,
                      uint32_t& maxRequests);
// Traces the direct rays of the volume's current requests ahead (k_direct, on
// the volume's stream) into its direct-term buffer and points p.directIn at
// it; false (p unchanged) when there are none to trace or on failure.
bool trace_direct_terms(DNvolume* vol, DNlightParams& p, uint32_t maxRequests);

// Kernel launchers (kernels.hip). All asynchronous on `stream`; they return a
// hipError_t value.
int dn_launch_draw(const DNdrawParams& p, hipStream_t stream);
int dn_launch_lighting(const DNlightParams& p, uint32_t maxRequests, hipStream_t stream);
// Direct rays of every voxel of the pass ahead of it, a wave holding 64 / N
// voxels x N directions (N = numDirectRays), into p.directOut.
int dn_launch_direct(const DNlightParams& p, uint32_t maxRequests, hipStream_t stream);
int dn_launch_lighting_epilogue(const DNlightParams& p, uint32_t tiles, const uint32_t* requests,
                                const uint32_t* numRequests, int copyBack, hipStream_t stream);
int dn_launch_visible(const DNvisibleParams& p, hipStream_t stream);
int dn_launch_upload(uint32_t* flags, uint32_t* lastUsed, DNchunkRec* chunks, uint4* voxels, uint32_t* fresh,
                     const DNuploadOp* ops, uint32_t numOps, const DNchunkRec* recs, const uint4* stagedVoxels,
                     hipStream_t stream);
int dn_launch_rays(const DNrayParams& p, hipStream_t stream);
int dn_launch_collect_requested(const uint32_t* flags, uint32_t tiles, uint32_t* list, uint32_t* count,
                                hipStream_t stream);
int dn_launch_mark_visible(uint32_t* flags, const uint32_t* list, uint32_t n, hipStream_t stream);
// Batched pass: unit prefix + counter reset, with `direct` the direct rays of
// the volumes whose parameters carry directOut (k_direct_batch), the persistent
// lighting kernel, then the per-volume epilogue and copy-back. `blockThreads` 64 or 256, `blocksPerCU`
// 0 = as many as fit. `maxTiles` / `maxRequests`: largest over the volumes.
int dn_launch_batch(const DNbatchParams& b, uint32_t blockThreads, uint32_t blocksPerCU, uint32_t maxTiles,
                    uint32_t maxRequests, bool epilogue, bool direct, hipStream_t stream);

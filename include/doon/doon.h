/*
 * doon.h -- C-ABI of the MI355X-native DoonEngine voxel pipeline.
 *
 * Part 1 (DN_*) is the reference's engine API, src/DoonEngine/voxel.h:156-371 and
 * src/DoonEngine/globals.h:11-81, with the same names, argument meaning and error
 * behaviour (bool false / NULL / -1 plus a message through g_DN_message_callback).
 * OpenGL types are replaced: GL buffer handles become device pointers, the output
 * texture becomes a caller-owned device RGBA32F buffer (DNtexture).
 *
 * Part 2 (DNX_*) are additions with no counterpart in the reference: bulk edits,
 * explicit residency, device-state export for parity checks, counters and
 * kernel timing, ray batches, batched multi-scene lighting and the radiance cache.
 *
 * Every entry point takes plain pointers and sizes; device pointers are HIP
 * device addresses (hipMalloc'd by the caller where the caller owns them).
 * Calls on one DNvolume are ordered on that volume's HIP stream; no call blocks
 * the host unless its comment says so.
 */
#ifndef DOON_MI355X_H
#define DOON_MI355X_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * globals.h:11-40 -- small vector structs
 */
typedef struct DNivec2 { int32_t x, y; } DNivec2;
typedef struct DNivec3 { int32_t x, y, z; } DNivec3;
typedef struct DNivec4 { int32_t x, y, z, w; } DNivec4;
typedef struct DNuvec2 { uint32_t x, y; } DNuvec2;
typedef struct DNuvec3 { uint32_t x, y, z; } DNuvec3;
typedef struct DNuvec4 { uint32_t x, y, z, w; } DNuvec4;

/* QuickMath types used by voxel.h (quickmath.h:111-175): plain floats, matrices
 * column-major m[column][row]. */
typedef struct DNvec2 { float x, y; } DNvec2;
typedef struct DNvec3 { float x, y, z; } DNvec3;
typedef struct DNvec4 { float x, y, z, w; } DNvec4;
typedef struct DNmat4 { float m[4][4]; } DNmat4;
typedef struct DNquaternion { float x, y, z, w; } DNquaternion;
/* QuickMath's DN_quaternion_identity (quickmath.h:1338-1348, QM_PREFIX DN_), which
 * the reference's callers pass as the orientation of the shape functions. */
static inline DNquaternion DN_quaternion_identity(void)
{
    DNquaternion q = {0.0f, 0.0f, 0.0f, 1.0f};
    return q;
}

/* ---------------------------------------------------------------------------
 * globals.h:45-63 -- message callback
 */
typedef enum DNmessageType {
    DN_MESSAGE_CPU_MEMORY,
    DN_MESSAGE_GPU_MEMORY,
    DN_MESSAGE_SHADER,
    DN_MESSAGE_FILE_IO
} DNmessageType;

typedef enum DNmessageSeverity {
    DN_MESSAGE_NOTE,
    DN_MESSAGE_ERROR,
    DN_MESSAGE_FATAL
} DNmessageSeverity;

/* globals.h:63. May be left NULL: messages of severity ERROR and FATAL then go
 * to stderr (the reference calls through a NULL pointer instead). */
extern void (*g_DN_message_callback)(DNmessageType, DNmessageSeverity, const char*);

/* ---------------------------------------------------------------------------
 * voxel.h:17-31 -- constants
 */
#define DN_CHUNK_SIZE 8
#define DN_CHUNK_LENGTH 512
#define DN_MAX_MATERIALS 256
#define DN_MATERIAL_EMPTY 255
#define DN_GAMMA 2.2f
#define DN_FLATTEN_INDEX(p, s) (p.x) + (s.x) * ((p.y) + (p.z) * (s.y))

/* ---------------------------------------------------------------------------
 * voxel.h:35-94 -- CPU-side data types (layouts identical to the reference)
 */
typedef struct DNcolor { uint8_t r, g, b; } DNcolor;

typedef struct DNvoxel {            /* voxel.h:41-46 */
    uint8_t material;
    DNvec3 normal;
    DNcolor albedo;
} DNvoxel;

typedef struct DNcompressedVoxel {  /* voxel.h:49-53 */
    uint32_t normal;                /* material << 24 | nx << 16 | ny << 8 | nz */
    uint32_t albedo;                /* r << 24 | g << 16 | b << 8 */
} DNcompressedVoxel;

typedef struct DNchunk {            /* voxel.h:56-64 */
    DNivec3 pos;
    bool updated;
    uint32_t numVoxels;
    uint32_t numVoxelsGpu;
    DNcompressedVoxel voxels[8][8][8];   /* indexed [x][y][z] */
} DNchunk;

typedef struct DNchunkHandle {      /* voxel.h:67-71 */
    uint8_t flag;                   /* 0 = does not exist, 1 = loaded on CPU */
    uint32_t chunkIndex;
} DNchunkHandle;

typedef struct DNvoxelNode {        /* voxel.h:74-79 */
    uint32_t size;
    size_t startPos;
    DNivec3 chunkPos;
} DNvoxelNode;

typedef struct DNmaterial {         /* voxel.h:82-94, 32 bytes, std140 */
    DNvec2 padding;                 /* padding.x > 0: emission strength of an emissive
                                       material (extension; 0 keeps the reference's 1.0) */
    uint32_t emissive;
    float opacity;
    float refractIndex;
    float specular;
    uint32_t reflectType;           /* 0 no sky, 1 sky, 2 highlight only */
    uint32_t shininess;
} DNmaterial;

/* Output image: caller-owned device buffer of width*height RGBA32F texels,
 * row-major, texel (x, y) at data[4*(y*width + x)]. Replaces `GLuint outputTexture`. */
typedef struct DNtexture {
    float* data;
    uint32_t width;
    uint32_t height;
} DNtexture;

struct DNgpuVolume;                 /* opaque device state of a volume */

typedef struct DNvolume {           /* voxel.h:97-142 */
    /* device buffers (replace glMapBufferID / glChunkBufferID / glVoxelBufferID) */
    void* gpuMap;                   /* READ ONLY  | per-tile flags words (u32) */
    void* gpuChunks;                /* READ ONLY  | per-tile 128-byte chunk records */
    void* gpuVoxels;                /* READ ONLY  | voxel pool, 16 bytes per voxel */

    DNuvec3 mapSize;                /* READ ONLY  | map size in chunks */
    size_t chunkCap;                /* READ ONLY  | CPU chunk capacity */
    size_t nextChunk;               /* READ ONLY  */
    size_t voxelCap;                /* READ ONLY  | GPU voxel pool capacity */
    size_t numVoxelNodes;           /* READ ONLY  | allocated pool nodes */
    size_t numLightingRequests;     /* READ ONLY  | refreshed by DNX_fetch_lighting_requests */
    size_t lightingRequestCap;      /* READ ONLY  */

    DNchunkHandle* map;             /* READ-WRITE */
    DNchunk* chunks;                /* READ-WRITE */
    DNmaterial* materials;          /* READ-WRITE | uploaded when changed (compared on DN_draw/DN_update_lighting) */
    uint32_t* lightingRequests;     /* READ ONLY  | host copy, filled by DNX_fetch_lighting_requests */
    DNvoxelNode* gpuVoxelLayout;    /* READ ONLY  | pool nodes (size, start, owner) */

    DNvec3 camPos;                  /* READ-WRITE | in chunks */
    DNvec3 camOrient;               /* READ-WRITE | {pitch, yaw, roll} degrees */
    float camFOV;                   /* READ-WRITE */
    uint32_t camViewMode;           /* READ-WRITE | 0 lighting .. 5 face normals */

    DNvec3 sunDir;                  /* READ-WRITE */
    DNvec3 sunStrength;             /* READ-WRITE */
    DNvec3 ambientLightStrength;    /* READ-WRITE */
    uint32_t diffuseBounceLimit;    /* READ-WRITE */
    uint32_t specBounceLimit;       /* READ-WRITE */
    float shadowSoftness;           /* READ-WRITE */

    bool useCubemap;                /* READ-WRITE | sample the sky from the cubemap given with DNX_set_sky_cubemap */
    uint32_t cubemapTex;            /* READ-WRITE | unused (a GL texture handle in the reference) */
    DNvec3 skyGradientBot;          /* READ-WRITE */
    DNvec3 skyGradientTop;          /* READ-WRITE */

    uint32_t frameNum;              /* READ ONLY  */
    float lastTime;                 /* READ ONLY  */

    /* extension fields */
    uint32_t numSpecularRays;       /* READ-WRITE | fixed specular rays per voxel; 15 = the reference's spherePoints */
    uint32_t numDirectRays;         /* READ-WRITE | 0 = one jittered shadow ray per sample (reference);
                                       N > 0 = N fixed, uniformly spaced shadow rays per sample */
    struct DNgpuVolume* gpu;        /* READ ONLY  | opaque */
} DNvolume;

typedef enum DNmemOp { DN_READ, DN_WRITE, DN_READ_WRITE } DNmemOp;   /* voxel.h:145-150 */

/* ---------------------------------------------------------------------------
 * voxel.h:156-371 -- the engine API
 */
bool DN_init(void);                                                    /* voxel.h:156 */
void DN_quit(void);                                                    /* voxel.h:158 */
DNvolume* DN_create_volume(DNuvec3 mapSize, unsigned int minChunks);   /* voxel.h:166 */
void DN_delete_volume(DNvolume* vol);                                  /* voxel.h:170 */
/* The .voxvol volume file (voxel.c:520-654): chunks with per-chunk palettes and
 * material runs, materials, camera, lighting, sky. Loading checks every field:
 * a truncated or malformed file gives an error message and NULL. */
DNvolume* DN_load_volume(const char* filePath, unsigned int minChunks); /* voxel.h:178 */
bool DN_save_volume(const char* filePath, DNvolume* vol);              /* voxel.h:184 */

void DN_set_view_projection_matrices(DNvolume* vol, float aspectRatio, float nearPlane, float farPlane,
                                     DNmat4* view, DNmat4* projection);                       /* voxel.h:197 */
/* voxel.h:207. rasterColorTexture / rasterDepthTexture must be -1 (no raster
 * compositing on this path); other values are reported and ignored. */
void DN_draw(DNvolume* vol, DNtexture outputTexture, DNmat4 view, DNmat4 projection,
             int rasterColorTexture, int rasterDepthTexture);
void DN_update_lighting(DNvolume* vol, int numDiffuseSamples, int maxDiffuseSamples, float time); /* voxel.h:215 */

int DN_add_chunk(DNvolume* vol, DNivec3 pos);                          /* voxel.h:225 */
void DN_remove_chunk(DNvolume* vol, DNivec3 pos);                      /* voxel.h:230 */
void DN_sync_gpu(DNvolume* vol, DNmemOp op, int lightingSplit);        /* voxel.h:237 */

bool DN_set_map_size(DNvolume* vol, DNuvec3 size);                     /* voxel.h:247 */
bool DN_set_max_chunks(DNvolume* vol, size_t num);                     /* voxel.h:254 */
bool DN_set_max_voxels_gpu(DNvolume* vol, size_t num);                 /* voxel.h:260 */
bool DN_set_max_lighting_requests(DNvolume* vol, size_t num);          /* voxel.h:267 */

bool DN_in_map_bounds(DNvolume* vol, DNivec3 pos);                     /* voxel.h:277 */
bool DN_in_chunk_bounds(DNivec3 pos);                                  /* voxel.h:282 */
DNvoxel DN_get_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos); /* voxel.h:290 */
DNcompressedVoxel DN_get_compressed_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos); /* voxel.h:297 */
void DN_set_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos, DNvoxel voxel);          /* voxel.h:305 */
void DN_set_compressed_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos, DNcompressedVoxel voxel); /* voxel.h:312 */
void DN_remove_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos);  /* voxel.h:318 */
bool DN_does_chunk_exist(DNvolume* vol, DNivec3 pos);                  /* voxel.h:325 */
bool DN_does_voxel_exist(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos);                  /* voxel.h:332 */
bool DN_step_map(DNvolume* vol, DNvec3 rayDir, DNvec3 rayPos, int maxSteps, DNivec3* hitPos,
                 DNvoxel* hitVoxel, DNivec3* hitNormal);               /* voxel.h:344 */

void DN_separate_position(DNivec3 pos, DNivec3* mapPos, DNivec3* chunkPos);  /* voxel.h:354 */
DNvec3 DN_cam_dir(DNvec3 orient);                                      /* voxel.h:360 */
DNcompressedVoxel DN_compress_voxel(DNvoxel voxel);                    /* voxel.h:366 */
DNvoxel DN_decompress_voxel(DNcompressedVoxel voxel);                  /* voxel.h:371 */

/* voxelShapes.h:19-28: the transform callback of the shape functions. Called
 * with the voxel's position (in voxels), its computed normal, the voxel given to
 * the shape (or, on a carve's rim, the voxel already there), the shape's
 * bounding box and the inverse of its transform; returns the voxel to place. */
typedef DNvoxel (*VoxelTransformFunc)(DNvec3 pos, DNvec3 normal, DNvoxel vox, DNvec3 shapeMin, DNvec3 shapeMax,
                                      DNmat4 invTransform, void* userData);

/* voxelShapes.h:40: sets every voxel within radius r of c to `voxel` (normals
 * point away from c), or carves them when voxel.material == DN_MATERIAL_EMPTY;
 * flipNormals re-points the rim's normals inwards. orient ({x, y, z, w}
 * quaternion) and userData only reach func; func NULL places `voxel` with the
 * computed normal. Runs on the host against the CPU chunks, like every edit;
 * DNX_shape_spheres is the batched device path for func == NULL. */
void DN_shape_sphere(DNvolume* vol, DNvoxel voxel, bool flipNormals, DNvec3 c, float r, DNquaternion orient,
                     VoxelTransformFunc func, void* userData);

/* ---------------------------------------------------------------------------
 * DNX_* -- extensions (no counterpart in the reference)
 */

/* Sizes of the public structs (0 DNvolume, 1 DNchunk, 2 DNmaterial, 3 DNchunkHandle,
 * 4 DNvoxelNode, 5 DNXcounters, 6 DNXtiming, 7 DNXrayHit, 8 DNXpixelHit, 9 DNvoxel;
 * 10/11 offsets of DNvolume.numSpecularRays / camPos, 12 DNXbatchConfig, 13 DNXrcConfig,
 * 14 DNXsphereEdit), for checking bindings. */
size_t DNX_struct_size(int which);

/* Messages and errors. */
void DNX_set_message_callback(void (*cb)(DNmessageType, DNmessageSeverity, const char*));
/* Number of ERROR/FATAL messages emitted so far (all volumes). */
uint32_t DNX_error_count(void);

/* Bulk edit: sets the voxels of a dense box of voxel coordinates [origin, origin+size)
 * (x fastest, then y, then z). normalWords/albedoWords use DNcompressedVoxel's
 * encoding; material 255 means empty. Same effect as DN_set_compressed_voxel per voxel. */
void DNX_set_voxels_dense(DNvolume* vol, DNivec3 origin, DNuvec3 size,
                          const uint32_t* normalWords, const uint32_t* albedoWords);
/* Sky cubemap for useCubemap (voxelShared.comp:291-292, main.c:335-353): six
 * square RGB8 faces of size x size texels in GL order (+X, -X, +Y, -Y, +Z, -Z),
 * each face's rows in glTexImage2D order (row 0 first), sampled like the
 * reference's GL_LINEAR / GL_CLAMP_TO_EDGE texture. NULL or size 0 removes it.
 * Replaces the reference's GL texture handle vol->glCubemapTex. */
bool DNX_set_sky_cubemap(DNvolume* vol, const uint8_t* faces, uint32_t size);

/* Batched sphere edits computed on the GPU. Applies the n edits in order with
 * exactly the effect of n DN_shape_sphere calls (voxelShapes.c:239-333, no
 * transform callback): the same voxel words, chunk-present flags, voxel
 * counts and `updated` flags, chunks added to the same pool slots
 * (DN_add_chunk order) and the same dirty-chunk order for the next sync. The
 * device keeps a copy of the CPU chunk words (4 KiB per map tile, allocated on
 * first use); tiles changed since through the edit API are re-sent on the next
 * call. Returns false (with a message) if a device step fails. */
typedef struct DNXsphereEdit {
    DNvoxel voxel;                  /* material DN_MATERIAL_EMPTY carves */
    DNvec3 center;
    float radius;
    uint32_t flipNormals;           /* carve only: turn the rim's normals inwards */
} DNXsphereEdit;
bool DNX_shape_spheres(DNvolume* vol, const DNXsphereEdit* edits, uint32_t n);

/* Reads the same dense box back (DN_get_compressed_voxel per voxel). Voxels of
 * tiles without a chunk, or outside the map, read as the cleared word pair
 * (normal 0xFFFFFFFF, albedo 0). */
void DNX_get_voxels_dense(DNvolume* vol, DNivec3 origin, DNuvec3 size, uint32_t* normalWords, uint32_t* albedoWords);

/* Residency: marks every chunk present on the CPU to be (re)loaded by the next
 * DN_sync_gpu, instead of waiting for rays to request it. */
void DNX_request_all_chunks(DNvolume* vol);

/* Bounds the GPU voxel pool at maxVoxels (rounded up to 512; 0 removes the
 * bound). A chunk to load takes a free pool node, else the node of the least
 * recently used chunk not entered by a ray in the last two syncs (voxel.c
 * :1554-1611), else the pool doubles up to this bound (and up to chunkCap *
 * 512 voxels, as the reference's resize), else its load waits for a later
 * request. Fails (false, message) when the pool is already larger. */
bool DNX_set_voxel_pool_limit(DNvolume* vol, size_t maxVoxels);
/* Chunks evicted to make room, and loads that had to wait, since the last reset. */
void DNX_get_residency_stats(DNvolume* vol, uint64_t* evictions, uint64_t* deferrals, bool reset);

/* Marks up to maxChunks exposed loaded chunks (map-index order) visible, as
 * primary rays hitting them would (voxelDraw.comp:118-121): the next DN_sync_gpu
 * requests lighting for all of them. A chunk is exposed when one of its voxels
 * has an empty, out-of-map or see-through face neighbour (computed on the host,
 * cached until the next edit). Returns the number of chunks marked. */
uint32_t DNX_mark_visible(DNvolume* vol, uint32_t maxChunks);

/* Device stream of the volume (a hipStream_t). */
void* DNX_get_stream(DNvolume* vol);
/* Blocks until all work queued on the volume's stream is done. Returns false on a device error. */
bool DNX_synchronize(DNvolume* vol);

/* Blocks; copies the current request list to vol->lightingRequests and returns its length. */
size_t DNX_fetch_lighting_requests(DNvolume* vol);

/* Device state in the reference's buffer layouts (voxelShared.comp:19-95), for
 * parity checks. Blocks. map: tiles*3 words (flags, lastUsed, voxelIndex);
 * chunks: tiles*24 words; voxels: voxelWords words (4 per voxel, pool order).
 * Any pointer may be NULL. Returns the pool size in voxels. */
size_t DNX_export_gpu_state(DNvolume* vol, uint32_t* map, uint32_t* chunks, uint32_t* voxels, size_t voxelWords);

/* Fixed direction tables used by the lighting pass (numSpecularRays / numDirectRays
 * points, 3 floats each). Returns the number of points written. */
uint32_t DNX_get_ray_points(DNvolume* vol, int which /* 0 specular, 1 direct */, float* out, uint32_t cap);

/* Ray counters accumulated by the kernels: step_map calls, chunk-grid trips,
 * voxel trips, hits, plus lit voxels and drawn pixels. */
typedef struct DNXcounters {
    uint64_t rays;          /* step_map calls */
    uint64_t mapSteps;      /* chunk-grid DDA trips */
    uint64_t voxelSteps;    /* voxel DDA trips */
    uint64_t hits;          /* step_map calls that hit */
    uint64_t litVoxels;     /* stored voxels written by lighting */
    uint64_t pixels;        /* pixels written by draw */
} DNXcounters;
/* Blocks; kind 0 = draw, 1 = lighting, 2 = ray batches. reset clears after reading. */
void DNX_get_counters(DNvolume* vol, int kind, DNXcounters* out, bool reset);

/* Kernel timing with HIP events recorded on the volume's stream around each
 * launch. Disabled by default. */
void DNX_enable_timing(DNvolume* vol, bool enable);
typedef struct DNXtiming {
    double drawMs, lightingMs, visibleMs, uploadMs, epilogueMs, raysMs;
    uint32_t drawLaunches, lightingLaunches, visibleLaunches, uploadLaunches, epilogueLaunches, raysLaunches;
} DNXtiming;
/* Blocks; returns accumulated times. reset clears after reading. */
void DNX_get_timing(DNvolume* vol, DNXtiming* out, bool reset);

/* One step_map call per ray on the device (voxelShared.comp:421). origins/dirs:
 * device arrays of 3*n floats in map units; out: device array of n DNXrayHit. */
typedef struct DNXrayHit {
    uint32_t hit, voxelIndex, mapSteps, voxelSteps;
    float pos[3];
    float normal[3];
} DNXrayHit;
void DNX_cast_rays(DNvolume* vol, uint32_t n, const float* origins, const float* dirs, bool ignoreFirst,
                   DNXrayHit* out);

/* Per-pixel hit record of DN_draw, for parity checks: device array of at least
 * width*height DNXpixelHit (capacityPixels entries), written while set; NULL disables. */
typedef struct DNXpixelHit {
    uint32_t voxelIndex, mapSteps, voxelSteps, pad;
    float face[3];
} DNXpixelHit;
void DNX_set_draw_hit_output(DNvolume* vol, DNXpixelHit* out, size_t capacityPixels);

/* Pre-quantisation lighting values of the last DN_update_lighting, for parity
 * checks: device array of 32 * requestCapacity slots of 6 floats (diffuse rgb,
 * spec rgb) and of pool indices (0xFFFFFFFF for empty lanes). */
void DNX_set_lighting_debug_output(DNvolume* vol, float* values, uint32_t* voxelIndices, size_t requestCapacity);

/* This pass's diffuse estimate of each voxel before it enters the running mean
 * (voxelLighting.comp:254-261 summed over the samples, divided by their count):
 * device array of 32 * requestCapacity float4 (rgb, 1 for a stored voxel and 0
 * for an empty lane). NULL disables. */
void DNX_set_lighting_sample_output(DNvolume* vol, float* samples, size_t requestCapacity);

/* Host-to-device traffic of DN_sync_gpu since the last reset: staged bytes
 * (upload ops, chunk records, voxels, occupancy bitmaps), chunks and voxels
 * uploaded. The reference re-uploads through glBufferSubData (voxel.c:1491-1695). */
void DNX_get_upload_stats(DNvolume* vol, uint64_t* bytes, uint64_t* chunks, uint64_t* voxels, bool reset);

/* Camera matrices for the draw kernel, from DN_draw's arguments: the inverse
 * view matrix and invCenteredView * invProjection (voxel.c:844-853). */
void DNX_draw_matrices(DNmat4 view, DNmat4 projection, DNmat4* invView, DNmat4* invViewProj);

/* Batched lighting: one persistent launch lights every volume of a batch, with
 * results bit-identical to calling DN_update_lighting on each volume in turn.
 * No counterpart in the reference (whose lighting pass is one glDispatchCompute
 * per volume, voxel.c:883-952). Each volume must have been DN_sync_gpu'd; the
 * batch orders itself after every volume's stream and each volume's stream
 * after the batch. The frame's launches are captured once into a hipGraph per
 * configuration and replayed. */
typedef struct DNXbatch DNXbatch;
typedef struct DNXbatchConfig {
    uint32_t requestsPerUnit;   /* requests one wave takes per queue grab: even, 2..64 */
    uint32_t blockThreads;      /* 64 or 256 */
    uint32_t blocksPerCU;       /* 0: as many as fit */
    uint32_t xcdQueues;         /* 1: one queue per XCD (8), volumes dealt round-robin; 0: one queue */
} DNXbatchConfig;
DNXbatch* DNX_batch_create(DNvolume* const* volumes, uint32_t count);
void DNX_batch_destroy(DNXbatch* batch);
bool DNX_batch_set_config(DNXbatch* batch, const DNXbatchConfig* config);
void DNX_batch_get_config(DNXbatch* batch, DNXbatchConfig* config);
/* Same arguments and semantics as DN_update_lighting, for every volume. */
bool DNX_batch_update_lighting(DNXbatch* batch, int numDiffuseSamples, int maxDiffuseSamples, float time);
/* Times every candidate configuration with compute-only passes (no stores) of
 * the current frame's requests, `repeats` launches each, keeps the fastest and
 * returns it. Volume state is left as it was. */
bool DNX_batch_autotune(DNXbatch* batch, int numDiffuseSamples, int maxDiffuseSamples, float time, int repeats,
                        DNXbatchConfig* best, float* bestMs);
/* Milliseconds of the batched launches (HIP events around each graph launch)
 * since the last reset, and their count. */
double DNX_batch_timing(DNXbatch* batch, uint32_t* launches, bool reset);

/* Learned radiance cache (no counterpart in the reference): a multiresolution
 * hash encoding of the voxel position (16 levels x 2 features, 2^log2Table
 * entries per level) plus the voxel normal and albedo feed an MLP of four
 * weight matrices (3 hidden layers of 64, ReLU, linear RGB output, no biases),
 * bf16 operands with fp32 accumulation on the matrix cores. It is trained
 * online with Adam on path-traced per-voxel diffuse samples drawn with
 * per-sample random numbers, and predicts each voxel's diffuse irradiance
 * (the quantity DN_update_lighting accumulates). With fixed direct rays
 * (numDirectRays > 0) the ambient and direct terms are deterministic: they are
 * evaluated exactly at prediction time and the network learns only the
 * path-traced indirect term. The plain accumulator keeps running beside it.
 * A sample row is 16 floats: position in voxel units (3), normal (3), albedo
 * (3), target (3), valid 0/1, and an exactly known term (3) that predictions
 * add to the network's output. */
typedef struct DNXradianceCache DNXradianceCache;
typedef struct DNXrcConfig {
    uint32_t log2Table;        /* 19 */
    uint32_t nMin, nMax;       /* coarsest and finest grid resolution: 16, 256 */
    float extent;              /* positions / extent fall in [0,1); 0: the longest map side in voxels */
    float lr, beta1, beta2, eps;   /* Adam: 1e-2, 0.9, 0.99, 1e-10 */
    uint32_t samplesPerFrame;  /* 65536 (multiple of 16) */
    uint32_t seed;             /* weight initialisation and sample draws */
    float emaDecay;            /* predictions use a debiased moving average of the parameters over the
                                  Adam steps with this decay (0.95); 0: the latest parameters */
    float lrDecaySteps;        /* step t (from 1) uses lr / (1 + (t - 1) / lrDecaySteps)^lrDecayPower; */
    float lrDecayPower;        /* lrDecaySteps 0 (default): constant lr. Default power 0.5 */
    float emaGrowth;           /* > 0 (default 8): step t's average decay is min(emaDecay, 1 - emaGrowth / t),
                                  a window of about t / emaGrowth steps while it grows; 0: emaDecay */
} DNXrcConfig;
void DNX_rc_default_config(DNXrcConfig* config);
DNXradianceCache* DNX_rc_create(DNvolume* vol, const DNXrcConfig* config /* NULL: defaults */);
void DNX_rc_destroy(DNXradianceCache* rc);
/* After DN_sync_gpu: takes `steps` Adam steps, each on samplesPerFrame fresh
 * samples path-traced on voxels drawn from the volume's current lighting requests. */
bool DNX_rc_train_frame(DNXradianceCache* rc, float time, int steps);
/* Predicted irradiance for every slot (32 per request) of the current request
 * list: device array of 32 * requestCapacity float4 (rgb, 1 for a stored voxel). */
bool DNX_rc_predict_requests(DNXradianceCache* rc, float* out, size_t requestCapacity);
/* Explicit samples (device rows, n a multiple of 16): gradients of the mean
 * squared-error loss accumulate until an Adam step (`step`) or DNX_rc_set_params. */
bool DNX_rc_train_samples(DNXradianceCache* rc, const float* samples, uint32_t n, bool step);
bool DNX_rc_predict_samples(DNXradianceCache* rc, const float* samples, uint32_t n, float* out /* device float4 */);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): DNX_rc_predict_samples with the same inputs and outputs
 * and the same results, but calling no hipMalloc, hipFree or allocating
 * routine. Its encoded input rows live in a fixed per-device array of
 * DNX_RC_NOALLOC_MAX_SAMPLES rows (24 MiB, held by each device until the
 * process ends), set up by DN_init: n above that maximum is rejected (false,
 * ERROR message). Every call on a device shares that array, so calls that
 * overlap on one device (other streams, other threads) race. */
#define DNX_RC_NOALLOC_MAX_SAMPLES 262144u
bool DNX_rc_predict_samples_noalloc(DNXradianceCache* rc, const float* samples, uint32_t n,
                                    float* out /* device float4 */);
/* Host copies of the fp32 master parameters (weights: 12288 floats, the padded
 * row-major W0 64x48, W1 64x64, W2 64x64, W3 16x64; table: 16 x 2^log2Table x 2).
 * set_params also clears the gradients and the Adam state. */
bool DNX_rc_set_params(DNXradianceCache* rc, const float* weights, const float* table);
bool DNX_rc_get_params(DNXradianceCache* rc, float* weights, float* table, float* weightGrad, float* tableGrad);
/* Reference estimates for scoring: adds samplesPerVoxel independent samples
 * (per-sample random numbers; exact ambient + direct with fixed direct rays) of
 * every stored voxel of the current request list into acc, a device array of 4
 * doubles per slot (rgb sums, sample count), in FP64. */
bool DNX_rc_reference(DNXradianceCache* rc, double* acc, size_t requestCapacity, uint32_t samplesPerVoxel,
                      uint32_t seed);
/* Device rows of the samples of the last DNX_rc_train_frame, and their count. */
const float* DNX_rc_last_samples(DNXradianceCache* rc, uint32_t* n);
/* dst[i] += (double)src[i] on the volume's stream (device arrays): FP64
 * accumulation of per-frame samples for reference estimates. */
bool DNX_accumulate_f64(DNvolume* vol, const float* src, double* dst, size_t n);

#ifdef __cplusplus
}
#endif

#endif

// Host engine: the reference's C API (src/DoonEngine/voxel.h:156-371) over HIP
// device state. CPU-side chunk storage and the edit API keep the reference's
// semantics (voxel.c:659-714, 1101-1316); the GPU side is designed for HIP:
// per-tile SoA device arrays, a buddy allocator for the voxel pool, one staged
// upload per sync, and device-side visible-chunk compaction.
#include "doon/doon.h"
#include "dn_internal.h"
#include "rc_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <set>
#include <thread>
#include <vector>

void (*g_DN_message_callback)(DNmessageType, DNmessageSeverity, const char*) = nullptr;

namespace {
std::atomic<uint32_t> g_errors{0};
}  // namespace

// Message dispatch through g_DN_message_callback (voxel.c:102); errors are
// also counted. Used by every host source of the library.
void dn_msg(DNmessageType type, DNmessageSeverity sev, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if(sev != DN_MESSAGE_NOTE) g_errors++;
    if(g_DN_message_callback)
        g_DN_message_callback(type, sev, buf);
    else if(sev != DN_MESSAGE_NOTE)
        fprintf(stderr, "doon: %s\n", buf);
}

namespace {

#define HIP_OK(expr, what)                                                                                  \
    ([&]() -> bool {                                                                                        \
        hipError_t e_ = (expr);                                                                             \
        if(e_ != hipSuccess)                                                                                \
        {                                                                                                   \
            dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "%s: %s", what, hipGetErrorString(e_));          \
            return false;                                                                                   \
        }                                                                                                   \
        return true;                                                                                        \
    }())

inline uint32_t material_of(uint32_t normalWord) { return normalWord >> 24; }

// ---------------------------------------------------------------------------
// buddy allocator over the voxel pool, in blocks of 16 << k voxels (k = 0..5)

struct PoolAlloc {
    std::set<uint32_t> freeSets[6];
    uint32_t cap = 0;   // voxels, multiple of 512

    static int cls(uint32_t n)
    {
        int k = 0;
        while((16u << k) < n) k++;
        return k;
    }
    void reset(uint32_t capacity)
    {
        for(auto& s : freeSets) s.clear();
        cap = 0;
        grow_to(capacity);
    }
    void grow_to(uint32_t capacity)
    {
        capacity = (capacity + 511u) & ~511u;
        for(uint32_t s = cap; s < capacity; s += 512u) freeSets[5].insert(s);
        if(capacity > cap) cap = capacity;
    }
    // returns start or UINT32_MAX
    uint32_t alloc(int k)
    {
        int j = k;
        while(j < 6 && freeSets[j].empty()) j++;
        if(j == 6) return UINT32_MAX;
        uint32_t start = *freeSets[j].begin();
        freeSets[j].erase(freeSets[j].begin());
        while(j > k)
        {
            j--;
            freeSets[j].insert(start + (16u << j));
        }
        return start;
    }
    void release(uint32_t start, int k)
    {
        while(k < 5)
        {
            uint32_t buddy = start ^ (16u << k);
            auto it = freeSets[k].find(buddy);
            if(it == freeSets[k].end()) break;
            freeSets[k].erase(it);
            start = std::min(start, buddy);
            k++;
        }
        freeSets[k].insert(start);
    }
};

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    int kind = 0;
};

}  // namespace

struct DNgpuVolume {
    hipStream_t stream = nullptr;
    uint32_t tiles = 0, occWords = 0, visBlocks = 0;

    uint32_t* dFlags = nullptr;
    uint32_t* dLastUsed = nullptr;
    DNchunkRec* dChunks = nullptr;
    uint4* dVoxels = nullptr;
    uint4* dVoxelsOut = nullptr;
    DNmat* dMats = nullptr;
    uint32_t* dOccLoaded = nullptr;
    uint32_t* dOccPresent = nullptr;
    uint32_t* dFresh = nullptr;
    uint32_t* dLitBits = nullptr;
    uint32_t* dVisProp = nullptr;
    uint32_t* dRequests = nullptr;
    uint32_t* dNumRequests = nullptr;
    uint32_t* dBlockSums = nullptr;
    bool presDirty = true;             // presLo/presHi and the column tables must be rebuilt from state
    int32_t presLo[3] = {0, 0, 0}, presHi[3] = {0, 0, 0};
    int32_t* dColTop = nullptr;        // DNsceneView::colTop, 4 x mapZ x mapX
    size_t colTopCap = 0;              // in entries
    bool colTopValid = false;          // dColTop holds the tables of the present state
    uint32_t* dWorkCounter = nullptr;  // [0] lighting requests taken, [1] collected requests, [2..3] draw quads,
                                       // [4] direct-ray requests taken
    DNdirectTerm* dDirect = nullptr;   // direct-ray terms per request slot (k_direct)
    size_t directCap = 0;              // in slots
    uint32_t* dReqList = nullptr;      // tiles requested by rays (collect)
    uint64_t* dCounters = nullptr;     // 3 kinds x CNT_SLOTS
    float* dPoints = nullptr;          // 2 x DN_MAX_POINTS x 3
    uint32_t specN = 0, directN = 0;
    std::vector<float> specPts, directPts;

    // staging: two pinned host buffers used alternately, one device buffer
    uint8_t* hStage[2] = {nullptr, nullptr};
    size_t hStageCap[2] = {0, 0};
    hipEvent_t stageDone[2] = {nullptr, nullptr};
    bool stageUsed[2] = {false, false};
    int stageNext = 0;
    uint8_t* dStage = nullptr;
    size_t dStageCap = 0;
    uint32_t* hSmall = nullptr;        // pinned readback words

    uchar4* dCube = nullptr;           // sky cubemap faces (DNX_set_sky_cubemap)
    uint32_t cubeSize = 0;

    // device sphere edits: device copy of the CPU chunk words (allocated on first use)
    uint2* dEditWords = nullptr;
    std::vector<uint8_t> editValid;    // per tile: device words equal the CPU chunk's
    std::vector<int32_t> editSlot;     // per tile: its index among a batch's tiles, -1 between calls
    uint8_t* hEdit = nullptr;          // pinned in/out staging
    size_t hEditCap = 0;
    uint8_t* dEdit = nullptr;          // device in/out staging
    size_t dEditCap = 0;
    hipEvent_t editEvents = nullptr;   // an edit batch's events are back on the host

    // host mirror of device tile state
    std::vector<uint8_t> state;        // 0 none, 1 unloaded, 2 loaded, 3 to load
    std::vector<uint32_t> nodeStart;
    std::vector<int8_t> nodeClass;     // -1 none
    std::vector<uint32_t> occLoaded, occPresent;
    std::vector<uint32_t> blocks;      // ceil(numVoxels/32) of loaded tiles
    bool occDirty = true;
    uint32_t numUnloaded = 0;          // tiles in state 1 or 3
    uint64_t loadedBlocks = 0;
    std::vector<uint32_t> dirty;
    std::vector<uint8_t> dirtyMark;
    bool fullScan = false;
    bool requestAll = false;
    bool resetPending = false;         // a failed upload: clear the device tiles and reload everything
    uint32_t poolLimit = 0;            // voxels; 0: the pool may grow to chunkCap * 512
    uint64_t evictions = 0, deferrals = 0;
    bool layoutDirty = true;
    // exposed chunks (a solid voxel with an empty or see-through face neighbour),
    // recomputed after edits, and the device list DNX_mark_visible uploads
    std::vector<uint8_t> exposed;
    bool exposedValid = false;
    uint32_t* dMarkList = nullptr;
    uint32_t markCount = 0, markMax = 0;
    bool markValid = false;
    PoolAlloc pool;

    DNmaterial lastMats[DN_MAX_MATERIALS];
    bool matsValid = false;
    bool anySpecular = false;

    bool timing = false;
    std::vector<Timer> pendingTimers;
    std::vector<Timer> freeTimers;
    double ms[6] = {0, 0, 0, 0, 0, 0};
    uint32_t launches[6] = {0, 0, 0, 0, 0, 0};

    DNXpixelHit* drawHitOut = nullptr;
    size_t drawHitCap = 0;
    float* dbgValues = nullptr;
    uint32_t* dbgIndex = nullptr;
    size_t dbgCap = 0;
    float* freshOut = nullptr;
    size_t freshCap = 0;
    uint64_t upBytes = 0, upChunks = 0, upVoxels = 0;   // host-to-device sync traffic
};

namespace {

enum { TK_DRAW = 0, TK_LIGHT, TK_VISIBLE, TK_UPLOAD, TK_EPILOGUE, TK_RAYS };

Timer timer_begin(DNgpuVolume* g, int kind)
{
    Timer t;
    if(!g->timing) return t;
    if(!g->freeTimers.empty()) { t = g->freeTimers.back(); g->freeTimers.pop_back(); }
    else { hipEventCreate(&t.a); hipEventCreate(&t.b); }
    t.kind = kind;
    hipEventRecord(t.a, g->stream);
    return t;
}
void timer_end(DNgpuVolume* g, Timer t)
{
    if(!t.a) return;
    hipEventRecord(t.b, g->stream);
    g->pendingTimers.push_back(t);
}
void timers_collect(DNgpuVolume* g)
{
    for(auto& t : g->pendingTimers)
    {
        hipEventSynchronize(t.b);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, t.a, t.b);
        g->ms[t.kind] += ms;
        g->launches[t.kind]++;
        g->freeTimers.push_back(t);
    }
    g->pendingTimers.clear();
}

template <typename T> bool dalloc(T** p, size_t count, const char* what)
{
    *p = nullptr;
    if(count == 0) count = 1;
    return HIP_OK(hipMalloc((void**)p, sizeof(T) * count), what);
}
template <typename T> void dfree(T*& p)
{
    if(p) hipFree(p);
    p = nullptr;
}

void free_tile_buffers(DNgpuVolume* g)
{
    dfree(g->dFlags); dfree(g->dLastUsed); dfree(g->dChunks); dfree(g->dOccLoaded); dfree(g->dOccPresent);
    dfree(g->dFresh); dfree(g->dLitBits); dfree(g->dVisProp); dfree(g->dRequests); dfree(g->dBlockSums);
    dfree(g->dReqList); dfree(g->dMarkList); dfree(g->dEditWords); dfree(g->dColTop);
    g->colTopCap = 0;
    g->colTopValid = false;
    g->editValid.clear();
    g->editSlot.clear();
    g->markValid = false;
    g->exposedValid = false;
}

bool alloc_tile_buffers(DNgpuVolume* g, uint32_t tiles)
{
    g->tiles = tiles;
    g->occWords = (tiles + 31u) / 32u;
    g->visBlocks = (tiles + 1023u) / 1024u;
    bool ok = dalloc(&g->dFlags, tiles, "map buffer") && dalloc(&g->dLastUsed, tiles, "map buffer") &&
              dalloc(&g->dChunks, tiles, "chunk buffer") && dalloc(&g->dOccLoaded, g->occWords, "occupancy") &&
              dalloc(&g->dOccPresent, g->occWords, "occupancy") && dalloc(&g->dFresh, g->occWords, "occupancy") &&
              dalloc(&g->dLitBits, g->occWords, "occupancy") && dalloc(&g->dVisProp, g->occWords, "occupancy") &&
              dalloc(&g->dRequests, (size_t)tiles * 16u, "lighting request buffer") &&
              dalloc(&g->dBlockSums, g->visBlocks, "request scan") && dalloc(&g->dReqList, tiles, "request list");
    if(!ok) return false;
    ok = HIP_OK(hipMemsetAsync(g->dFlags, 0, sizeof(uint32_t) * tiles, g->stream), "map buffer clear") &&
         HIP_OK(hipMemsetAsync(g->dLastUsed, 0, sizeof(uint32_t) * tiles, g->stream), "map buffer clear") &&
         HIP_OK(hipMemsetAsync(g->dChunks, 0, sizeof(DNchunkRec) * tiles, g->stream), "chunk buffer clear");
    for(uint32_t* b : {g->dOccLoaded, g->dOccPresent, g->dFresh, g->dLitBits, g->dVisProp})
        ok = ok && HIP_OK(hipMemsetAsync(b, 0, sizeof(uint32_t) * g->occWords, g->stream), "occupancy clear");
    if(!ok) return false;
    g->state.assign(tiles, 0);
    g->presDirty = true;
    g->nodeStart.assign(tiles, 0);
    g->nodeClass.assign(tiles, -1);
    g->occLoaded.assign(g->occWords, 0);
    g->occPresent.assign(g->occWords, 0);
    g->blocks.assign(tiles, 0);
    g->dirtyMark.assign(tiles, 0);
    g->dirty.clear();
    g->numUnloaded = 0;
    g->loadedBlocks = 0;
    g->occDirty = true;
    return true;
}

bool grow_pool(DNvolume* vol, uint32_t newCap)
{
    DNgpuVolume* g = vol->gpu;
    newCap = (newCap + 511u) & ~511u;
    if(newCap <= g->pool.cap) return true;
    uint4* nv = nullptr;
    if(!dalloc(&nv, (size_t)newCap, "voxel buffer")) return false;
    const size_t old = g->dVoxels ? g->pool.cap : 0;
    // free slots read as zero words, so exported pools are deterministic
    if((old && !HIP_OK(hipMemcpyAsync(nv, g->dVoxels, sizeof(uint4) * old, hipMemcpyDeviceToDevice, g->stream),
                       "voxel buffer copy")) ||
       !HIP_OK(hipMemsetAsync(nv + old, 0, sizeof(uint4) * (newCap - old), g->stream), "voxel buffer clear") ||
       !HIP_OK(hipStreamSynchronize(g->stream), "voxel buffer resize"))
    {
        hipStreamSynchronize(g->stream);
        dfree(nv);
        return false;
    }
    dfree(g->dVoxels);
    g->dVoxels = nv;
    if(g->dVoxelsOut)
    {
        dfree(g->dVoxelsOut);
        if(!dalloc(&g->dVoxelsOut, (size_t)newCap, "voxel output buffer")) return false;
    }
    g->pool.grow_to(newCap);
    vol->voxelCap = g->pool.cap;
    vol->gpuVoxels = g->dVoxels;
    return true;
}

void mark_dirty(DNgpuVolume* g, uint32_t t)
{
    if(t >= g->tiles) return;
    if(!g->editValid.empty()) g->editValid[t] = 0;
    g->exposedValid = false;
    g->markValid = false;
    if(!g->dirtyMark[t]) { g->dirtyMark[t] = 1; g->dirty.push_back(t); }
}

inline void set_bit(std::vector<uint32_t>& v, uint32_t t, bool on)
{
    if(on) v[t >> 5] |= 1u << (t & 31u);
    else v[t >> 5] &= ~(1u << (t & 31u));
}

void set_state(DNgpuVolume* g, uint32_t t, uint8_t st)
{
    uint8_t old = g->state[t];
    if(old == st) return;
    if((old != 0) != (st != 0)) g->presDirty = true;
    if(old == 1 || old == 3) g->numUnloaded--;
    if(st == 1 || st == 3) g->numUnloaded++;
    g->state[t] = st;
    set_bit(g->occLoaded, t, st == 2);
    set_bit(g->occPresent, t, st == 1 || st == 3);
    g->occDirty = true;
}

void free_node(DNgpuVolume* g, uint32_t t)
{
    if(g->nodeClass[t] >= 0)
    {
        g->pool.release(g->nodeStart[t], g->nodeClass[t]);
        g->nodeClass[t] = -1;
        g->layoutDirty = true;
    }
    g->loadedBlocks -= g->blocks[t];
    g->blocks[t] = 0;
}

// _DN_check_face_visible / _DN_chunk_to_gpu (voxel.c:1391-1461): per-chunk face
// culling, occupancy mask, rank prefix, albedo linearisation.
uint8_t g_linLUT[256];
bool g_lutReady = false;

void build_lut()
{
    if(g_lutReady) return;
    for(int v = 0; v < 256; v++)
    {
        float x = (float)v * 0.00392156862f;
        x = powf(x, DN_GAMMA) * 255.0f;
        g_linLUT[v] = (uint8_t)x;
    }
    g_lutReady = true;
}

uint32_t chunk_to_gpu(const DNmaterial* mats, const DNchunk& ch, DNchunkRec& rec, uint4* vox)
{
    memset(&rec, 0, sizeof(rec));
    rec.pos[0] = ch.pos.x; rec.pos[1] = ch.pos.y; rec.pos[2] = ch.pos.z;
    auto faceVisible = [&](int x, int y, int z) -> bool {
        if(x < 0 || y < 0 || z < 0 || x >= 8 || y >= 8 || z >= 8) return true;
        uint32_t m = material_of(ch.voxels[x][y][z].normal);
        return m == DN_MATERIAL_EMPTY || mats[m].opacity < 1.0f;
    };
    uint32_t n = 0;
    for(int z = 0; z < 8; z++)
    for(int y = 0; y < 8; y++)
    for(int x = 0; x < 8; x++)
    {
        uint32_t index = (uint32_t)(x + 8 * (y + 8 * z));
        if((index & 127u) == 0 && index != 0) rec.partial[(index >> 7) - 1] = n;
        const DNcompressedVoxel& v = ch.voxels[x][y][z];
        if(material_of(v.normal) == DN_MATERIAL_EMPTY) continue;
        if(!(faceVisible(x + 1, y, z) || faceVisible(x - 1, y, z) || faceVisible(x, y + 1, z) ||
             faceVisible(x, y - 1, z) || faceVisible(x, y, z + 1) || faceVisible(x, y, z - 1)))
            continue;
        rec.mask[index >> 5] |= 1u << (index & 31u);
        uint32_t a = v.albedo;
        uint32_t r = g_linLUT[(a >> 24) & 255u], gg = g_linLUT[(a >> 16) & 255u], b = g_linLUT[(a >> 8) & 255u];
        vox[n] = make_uint4(v.normal, (r << 24) | (gg << 16) | (b << 8), 0u, 0u);
        n++;
    }
    rec.numVoxels = n;
    rec.samples = 0;
    return n;
}

bool ensure_stage(DNgpuVolume* g, int slot, size_t bytes)
{
    if(g->hStageCap[slot] < bytes)
    {
        if(g->hStage[slot])
        {
            hipEventSynchronize(g->stageDone[slot]);
            hipHostFree(g->hStage[slot]);
            g->hStage[slot] = nullptr;
            g->hStageCap[slot] = 0;
        }
        size_t cap = std::max(bytes, (size_t)1 << 20);
        if(!HIP_OK(hipHostMalloc((void**)&g->hStage[slot], cap, hipHostMallocDefault), "pinned staging"))
        {
            g->hStageCap[slot] = 0;
            return false;
        }
        g->hStageCap[slot] = cap;
    }
    if(g->dStageCap < bytes)
    {
        hipStreamSynchronize(g->stream);
        dfree(g->dStage);
        size_t cap = std::max(bytes, (size_t)1 << 20);
        if(!dalloc(&g->dStage, cap, "device staging")) { g->dStageCap = 0; return false; }
        g->dStageCap = cap;
    }
    return true;
}

int acquire_stage_slot(DNgpuVolume* g)
{
    int s = g->stageNext;
    g->stageNext ^= 1;
    if(g->stageUsed[s]) hipEventSynchronize(g->stageDone[s]);
    return s;
}

void release_stage_slot(DNgpuVolume* g, int s)
{
    hipEventRecord(g->stageDone[s], g->stream);
    g->stageUsed[s] = true;
}

bool upload_materials(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    if(g->matsValid && memcmp(g->lastMats, vol->materials, sizeof(g->lastMats)) == 0) return true;
    memcpy(g->lastMats, vol->materials, sizeof(g->lastMats));
    g->matsValid = true;
    g->anySpecular = false;
    for(int i = 0; i < DN_MAX_MATERIALS; i++)
        if(vol->materials[i].specular > 0.0f && vol->materials[i].reflectType <= 1u) g->anySpecular = true;
    int s = acquire_stage_slot(g);
    if(!ensure_stage(g, s, sizeof(g->lastMats))) return false;
    memcpy(g->hStage[s], g->lastMats, sizeof(g->lastMats));
    bool ok = HIP_OK(hipMemcpyAsync(g->dMats, g->hStage[s], sizeof(g->lastMats), hipMemcpyHostToDevice, g->stream),
                     "material upload");
    release_stage_slot(g, s);
    return ok;
}

// Fixed direction sets on the unit sphere. N = 15 is the reference's
// spherePoints table (voxelLighting.comp:23), kept value for value so the
// reference's specular rays are reproduced exactly; other N use the golden-angle
// (Fibonacci) construction the table approximates: y = 1 - 2i/(N-1),
// r = sqrt(1 - y^2), phi = i * pi * (3 - sqrt 5), point (r cos phi, y, r sin phi).
const float kSpherePoints15[15][3] = {
    {0.000000f, 1.000000f, 0.000000f},    {-0.379803f, 0.857143f, 0.347931f},  {0.061185f, 0.714286f, -0.697174f},
    {0.499316f, 0.571429f, 0.651270f},    {-0.889696f, 0.428571f, -0.157375f}, {0.808584f, 0.285714f, -0.514354f},
    {-0.256942f, 0.142857f, 0.955810f},   {-0.460906f, 0.000000f, -0.887449f}, {0.929687f, -0.142857f, 0.339521f},
    {-0.885815f, -0.285714f, 0.365650f},  {0.382949f, -0.428571f, -0.818338f}, {0.245607f, -0.571429f, 0.783037f},
    {-0.605521f, -0.714286f, -0.350913f}, {0.503065f, -0.857143f, -0.110596f}, {-0.000000f, -1.000000f, 0.000000f}};

std::vector<float> fibonacci_points(uint32_t n)
{
    std::vector<float> out(3 * (size_t)n, 0.0f);
    if(n == 15)
    {
        for(int i = 0; i < 15; i++)
            for(int k = 0; k < 3; k++) out[3 * i + k] = kSpherePoints15[i][k];
        return out;
    }
    const double golden = 3.14159265358979323846 * (3.0 - std::sqrt(5.0));
    for(uint32_t i = 0; i < n; i++)
    {
        double y = n > 1 ? 1.0 - 2.0 * (double)i / (double)(n - 1) : 1.0;
        double r = std::sqrt(std::max(0.0, 1.0 - y * y));
        double phi = golden * (double)i;
        out[3 * i + 0] = (float)(std::cos(phi) * r);
        out[3 * i + 1] = (float)y;
        out[3 * i + 2] = (float)(std::sin(phi) * r);
    }
    return out;
}

bool ensure_points(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    uint32_t ns = vol->numSpecularRays, nd = vol->numDirectRays;
    if(ns < 1 || ns > DN_MAX_POINTS || nd > DN_MAX_POINTS)
    {
        dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR, "numSpecularRays must be in [1, %d] and numDirectRays in [0, %d]",
               DN_MAX_POINTS, DN_MAX_POINTS);
        return false;
    }
    if(ns == g->specN && nd == g->directN) return true;
    g->specPts = fibonacci_points(ns);
    g->directPts = fibonacci_points(nd);
    std::vector<float> all(2 * DN_MAX_POINTS * 3, 0.0f);
    std::copy(g->specPts.begin(), g->specPts.end(), all.begin());
    std::copy(g->directPts.begin(), g->directPts.end(), all.begin() + DN_MAX_POINTS * 3);
    hipStreamSynchronize(g->stream);
    if(!HIP_OK(hipMemcpy(g->dPoints, all.data(), sizeof(float) * all.size(), hipMemcpyHostToDevice), "ray table"))
        return false;
    g->specN = ns;
    g->directN = nd;
    return true;
}

// The reference's hash RNG (voxelLighting.comp:29-57), on the host in FP32.
inline float frand(float seed)
{
    float v = sinf(seed) * 43758.5453f;
    return (v - floorf(v)) * 2.0f - 1.0f;
}
void rand_unit_sphere(float seed, float out[3])
{
    float p[3] = {0, 0, 0};
    for(int k = 0; k < 256; k++)
    {
        p[0] = frand(seed); p[1] = frand(seed * 2.0f); p[2] = frand(seed * 3.0f);
        seed = seed + 1.0f;
        if(p[0] * p[0] + p[1] * p[1] + p[2] * p[2] >= 1.0f) continue;
        break;
    }
    out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
}

void frame_random(float time, uint32_t nds, uint32_t dbl, DNframeRandom& r)
{
    memset(&r, 0, sizeof(r));
    for(uint32_t i = 0; i < nds && i < DN_MAX_SAMPLES; i++)
    {
        float seed = time * (float)(i + 1);
        for(uint32_t j = 0; j < dbl && j < DN_MAX_BOUNCES; j++)
        {
            rand_unit_sphere(seed + (float)j, r.diffuseDir[i][j]);
            r.glossyPick[i][j] = frand(seed + (float)dbl + (float)j);
        }
        rand_unit_sphere(time * (float)(i + 1 + nds), r.shadowDir[i]);
    }
    for(uint32_t j = 0; j < dbl && j < DN_MAX_BOUNCES; j++) rand_unit_sphere(time + (float)j, r.glossyDir[j]);
}

// DNsceneView::colTop from the tile states: per column the highest present
// tile, then for each quadrant its maximum over the columns at or beyond a
// column in the quadrant's x and z directions. Uploaded through a staging
// slot on the volume's stream; false (an ERROR reported) leaves the rule off.
bool upload_column_tops(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    const uint32_t X = vol->mapSize.x, Y = vol->mapSize.y, Z = vol->mapSize.z;
    const size_t plane = (size_t)X * Z, n = 4 * plane;
    if(g->colTopCap < n)
    {
        hipStreamSynchronize(g->stream);
        dfree(g->dColTop);
        g->colTopCap = 0;
        if(!dalloc(&g->dColTop, n, "column table")) return false;
        g->colTopCap = n;
    }
    const int s = acquire_stage_slot(g);
    if(!ensure_stage(g, s, n * sizeof(int32_t))) return false;
    int32_t* t = (int32_t*)g->hStage[s];
    int32_t* top = t;                               // quadrant 0 first holds the column tops
    std::fill(top, top + plane, -1);
    for(uint32_t z = 0; z < Z; z++)
        for(uint32_t y = 0; y < Y; y++)
            for(uint32_t x = 0; x < X; x++)
                if(g->state[x + (size_t)X * (y + (size_t)Y * z)] != 0) top[(size_t)z * X + x] = (int32_t)y;
    for(int q = 3; q >= 0; q--)                     // quadrant 0 last: it reads the tops in place
    {
        int32_t* o = t + (size_t)q * plane;
        if(q != 0) std::copy(top, top + plane, o);
        const bool up_x = (q & 1) != 0, up_z = (q & 2) != 0;
        for(uint32_t z = 0; z < Z; z++)             // along x within each row
        {
            int32_t* row = o + (size_t)z * X;
            if(up_x)
            {
                for(uint32_t x = X - 1; x-- > 0;) row[x] = std::max(row[x], row[x + 1]);
            }
            else
            {
                for(uint32_t x = 1; x < X; x++) row[x] = std::max(row[x], row[x - 1]);
            }
        }
        for(uint32_t k = 1; k < Z; k++)             // then along z
        {
            const uint32_t z = up_z ? Z - 1 - k : k, from = up_z ? z + 1 : z - 1;
            for(uint32_t x = 0; x < X; x++)
                o[(size_t)z * X + x] = std::max(o[(size_t)z * X + x], o[(size_t)from * X + x]);
        }
    }
    const bool ok = HIP_OK(hipMemcpyAsync(g->dColTop, t, n * sizeof(int32_t), hipMemcpyHostToDevice, g->stream),
                           "column table upload");
    release_stage_slot(g, s);
    return ok;
}

DNsceneView scene_view(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    DNsceneView s;
    s.occLoaded = g->dOccLoaded;
    s.occPresent = g->dOccPresent;
    s.flags = g->dFlags;
    s.lastUsed = g->dLastUsed;
    s.chunks = g->dChunks;
    s.voxels = g->dVoxels;
    s.mats = g->dMats;
    s.mapX = vol->mapSize.x; s.mapY = vol->mapSize.y; s.mapZ = vol->mapSize.z;
    s.hasUnloaded = g->numUnloaded > 0 ? 1u : 0u;
    s.sunStrength[0] = vol->sunStrength.x; s.sunStrength[1] = vol->sunStrength.y; s.sunStrength[2] = vol->sunStrength.z;
    s.skyBot[0] = vol->skyGradientBot.x; s.skyBot[1] = vol->skyGradientBot.y; s.skyBot[2] = vol->skyGradientBot.z;
    s.skyTop[0] = vol->skyGradientTop.x; s.skyTop[1] = vol->skyGradientTop.y; s.skyTop[2] = vol->skyGradientTop.z;
    s.cube = (vol->useCubemap && g->dCube) ? g->dCube : nullptr;
    s.cubeSize = g->cubeSize;
    if(g->presDirty)
    {
        // box of the present tiles, rebuilt after a sync changed which tiles are present
        int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
        const uint32_t X = vol->mapSize.x, Y = vol->mapSize.y;
        for(uint32_t t = 0; t < g->tiles; t++)
            if(g->state[t] != 0)
            {
                const int32_t c[3] = {(int32_t)(t % X), (int32_t)((t / X) % Y), (int32_t)(t / (X * Y))};
                for(int a = 0; a < 3; a++)
                {
                    lo[a] = std::min(lo[a], c[a]);
                    hi[a] = std::max(hi[a], c[a]);
                }
            }
        for(int a = 0; a < 3; a++) { g->presLo[a] = lo[a]; g->presHi[a] = hi[a]; }
        g->colTopValid = upload_column_tops(vol);
        g->presDirty = false;
    }
    for(int a = 0; a < 3; a++) { s.presLo[a] = g->presLo[a]; s.presHi[a] = g->presHi[a]; }
    s.colTop = g->colTopValid ? g->dColTop : nullptr;
    return s;
}

// Direct-ray term buffer (k_direct) for up to maxReq requests. Growing frees
// the old buffer after the device is idle (a batch graph on another stream
// may still read it).
bool ensure_direct(DNgpuVolume* g, uint32_t maxReq)
{
    const size_t need = (size_t)maxReq * 32u;
    if(g->dDirect && g->directCap >= need) return true;
    hipDeviceSynchronize();
    dfree(g->dDirect);
    g->directCap = 0;
    if(!dalloc(&g->dDirect, need, "direct-ray terms")) return false;
    g->directCap = need;
    return true;
}

void clear_chunk_voxels(DNvolume* vol, size_t index)
{
    DNchunk& c = vol->chunks[index];
    for(int x = 0; x < 8; x++)
    for(int y = 0; y < 8; y++)
    for(int z = 0; z < 8; z++)
    {
        c.voxels[x][y][z].normal = UINT32_MAX;
        c.voxels[x][y][z].albedo = 0;
    }
}

void clear_chunk_header(DNvolume* vol, size_t index)
{
    DNchunk& c = vol->chunks[index];
    c.pos = DNivec3{-1, -1, -1};
    c.updated = false;
    c.numVoxels = 0;
    c.numVoxelsGpu = 0;
}

void clear_chunk(DNvolume* vol, size_t index)   // voxel.c:1353-1363
{
    clear_chunk_header(vol, index);
    clear_chunk_voxels(vol, index);
}

inline uint32_t flat(DNvolume* vol, DNivec3 p)
{
    return (uint32_t)p.x + vol->mapSize.x * ((uint32_t)p.y + (uint32_t)p.z * vol->mapSize.y);
}

void rebuild_layout(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    if(!g->layoutDirty) return;
    size_t n = 0;
    for(uint32_t t = 0; t < g->tiles; t++) n += g->nodeClass[t] >= 0;
    DNvoxelNode* nl = (DNvoxelNode*)realloc(vol->gpuVoxelLayout, sizeof(DNvoxelNode) * (n ? n : 1));
    if(!nl) return;
    vol->gpuVoxelLayout = nl;
    size_t k = 0;
    for(uint32_t t = 0; t < g->tiles; t++)
    {
        if(g->nodeClass[t] < 0) continue;
        nl[k].size = 16u << g->nodeClass[t];
        nl[k].startPos = g->nodeStart[t];
        uint32_t x = t % vol->mapSize.x, y = (t / vol->mapSize.x) % vol->mapSize.y, z = t / (vol->mapSize.x * vol->mapSize.y);
        nl[k].chunkPos = DNivec3{(int32_t)x, (int32_t)y, (int32_t)z};
        k++;
    }
    vol->numVoxelNodes = n;
    g->layoutDirty = false;
}

// Ray-requested tiles (flag 3 written by step_map, voxelShared.comp:462-466).
// Blocks; only runs while some tile is unloaded.
// Returns false (after an ERROR message) when a device step fails; the
// requests then stay flagged on the device and the next sync collects them.
bool collect_ray_requests(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    // after every sync each tile is 0, 1 or 2, so numUnloaded counts state-1 tiles
    if(g->numUnloaded == 0) return true;
    uint32_t* dCount = g->dWorkCounter + 1;
    if(!HIP_OK(hipMemsetAsync(dCount, 0, sizeof(uint32_t), g->stream), "ray request count")) return false;
    int e = dn_launch_collect_requested(g->dFlags, g->tiles, g->dReqList, dCount, g->stream);
    if(e)
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "ray request collection launch failed (%d)", e);
        return false;
    }
    if(!HIP_OK(hipMemcpyAsync(g->hSmall, dCount, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream),
               "ray request count readback") ||
       !HIP_OK(hipStreamSynchronize(g->stream), "ray request collection"))
        return false;
    uint32_t n = std::min(g->hSmall[0], g->tiles);
    if(n == 0) return true;
    std::vector<uint32_t> list(n);
    if(!HIP_OK(hipMemcpy(list.data(), g->dReqList, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "ray request readback"))
        return false;
    for(uint32_t t : list)
        if(t < g->tiles && g->state[t] == 1) { set_state(g, t, 3); mark_dirty(g, t); }
    return true;
}

// After a failed upload the device's tile state can no longer be trusted to
// match the host mirror: the next sync clears every tile's device flags and
// reloads every CPU chunk from scratch.
void schedule_device_reset(DNgpuVolume* g)
{
    g->resetPending = true;
}

// Least-recently-used pool node of class >= k among loaded chunks not entered
// by a ray in the last two syncs (voxel.c:1554-1604: lastUsed > 1, oldest
// first; among equal ages the smaller node, then the lower map index). Node
// lists are built once per sync from the device's lastUsed words.
struct EvictionQueue {
    std::vector<std::pair<uint32_t, uint32_t>> byClass[6];   // (age, tile), oldest last
    size_t head[6] = {0, 0, 0, 0, 0, 0};
    bool ready = false;
};

bool build_eviction_queue(DNvolume* vol, EvictionQueue& q)
{
    DNgpuVolume* g = vol->gpu;
    q.ready = true;
    std::vector<uint32_t> ages(g->tiles);
    if(!HIP_OK(hipMemcpyAsync(ages.data(), g->dLastUsed, sizeof(uint32_t) * g->tiles, hipMemcpyDeviceToHost, g->stream),
               "lastUsed readback") ||
       !HIP_OK(hipStreamSynchronize(g->stream), "lastUsed readback"))
        return false;
    for(uint32_t t = 0; t < g->tiles; t++)
        if(g->state[t] == 2 && g->nodeClass[t] >= 0 && g->dirtyMark[t] != 2 && ages[t] > 1u)
            q.byClass[g->nodeClass[t]].push_back({ages[t], t});
    for(auto& v : q.byClass)
        std::sort(v.begin(), v.end(), [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
            return a.first != b.first ? a.first > b.first : a.second < b.second;
        });
    return true;
}

// Pops the chunk to evict for a node of class k, or UINT32_MAX.
uint32_t pop_eviction(DNgpuVolume* g, EvictionQueue& q, int k)
{
    int best = -1;
    for(int j = k; j < 6; j++)
    {
        auto& v = q.byClass[j];
        // skip chunks unloaded since, or loaded by this sync (dirtyMark 2)
        while(q.head[j] < v.size() && (g->state[v[q.head[j]].second] != 2 || g->dirtyMark[v[q.head[j]].second] == 2))
            q.head[j]++;
        if(q.head[j] >= v.size()) continue;
        if(best < 0 || v[q.head[j]].first > q.byClass[best][q.head[best]].first) best = j;
    }
    if(best < 0) return UINT32_MAX;
    return q.byClass[best][q.head[best]++].second;
}

}  // namespace

// ===========================================================================
// INITIALIZATION (voxel.c:119-295)

extern "C" bool DN_init(void)
{
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess || n < 1)
    {
        dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_FATAL, "no HIP device available");
        return false;
    }
    build_lut();
    // every device's address of the allocation-free prediction's scratch array
    if(!dn_noalloc_init())
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "allocation-free prediction: scratch address lookup failed");
    return true;
}

extern "C" void DN_quit(void) {}

extern "C" DNvolume* DN_create_volume(DNuvec3 mapSize, unsigned int minChunks)
{
    build_lut();
    if(mapSize.x == 0 || mapSize.y == 0 || mapSize.z == 0 ||
       (uint64_t)mapSize.x * mapSize.y * mapSize.z >= (1ull << 28))
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_FATAL, "invalid map size");
        return nullptr;
    }
    DNvolume* vol = (DNvolume*)calloc(1, sizeof(DNvolume));
    if(!vol) { dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_FATAL, "failed to allocate volume"); return nullptr; }
    DNgpuVolume* g = new DNgpuVolume();
    vol->gpu = g;
    const uint32_t tiles = mapSize.x * mapSize.y * mapSize.z;
    if(!HIP_OK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "stream") ||
       !alloc_tile_buffers(g, tiles) || !dalloc(&g->dMats, DN_MAX_MATERIALS, "material buffer") ||
       !dalloc(&g->dNumRequests, 1, "request count") || !dalloc(&g->dWorkCounter, 6, "work counter") ||
       !dalloc(&g->dCounters, 3 * CNT_SLOTS, "counters") || !dalloc(&g->dPoints, 2 * DN_MAX_POINTS * 3, "ray table") ||
       !HIP_OK(hipHostMalloc((void**)&g->hSmall, 64, hipHostMallocDefault), "pinned readback") ||
       !HIP_OK(hipEventCreate(&g->stageDone[0]), "event") || !HIP_OK(hipEventCreate(&g->stageDone[1]), "event"))
    {
        DN_delete_volume(vol);
        return nullptr;
    }
    if(!HIP_OK(hipMemsetAsync(g->dNumRequests, 0, sizeof(uint32_t), g->stream), "request count") ||
       !HIP_OK(hipMemsetAsync(g->dWorkCounter, 0, 6 * sizeof(uint32_t), g->stream), "work counter") ||
       !HIP_OK(hipMemsetAsync(g->dCounters, 0, 3 * CNT_SLOTS * sizeof(uint64_t), g->stream), "counters") ||
       !HIP_OK(hipMemsetAsync(g->dMats, 0, sizeof(DNmat) * DN_MAX_MATERIALS, g->stream), "material buffer"))
    {
        DN_delete_volume(vol);
        return nullptr;
    }

    size_t numChunks = std::min<size_t>(tiles, minChunks);
    if(numChunks < 1) numChunks = 1;
    uint32_t voxelCap = (uint32_t)std::max<size_t>(512, DN_CHUNK_LENGTH * numChunks / 2);
    g->pool.reset(0);
    if(!grow_pool(vol, voxelCap)) { DN_delete_volume(vol); return nullptr; }

    vol->map = (DNchunkHandle*)calloc(tiles, sizeof(DNchunkHandle));
    vol->chunks = (DNchunk*)malloc(sizeof(DNchunk) * numChunks);
    vol->materials = (DNmaterial*)calloc(DN_MAX_MATERIALS, sizeof(DNmaterial));
    vol->lightingRequests = (uint32_t*)malloc(sizeof(uint32_t) * numChunks);
    vol->gpuVoxelLayout = (DNvoxelNode*)malloc(sizeof(DNvoxelNode));
    if(!vol->map || !vol->chunks || !vol->materials || !vol->lightingRequests || !vol->gpuVoxelLayout)
    {
        dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_FATAL, "failed to allocate memory for volume");
        DN_delete_volume(vol);
        return nullptr;
    }
    vol->mapSize = mapSize;
    for(size_t i = 0; i < numChunks; i++) clear_chunk(vol, i);
    vol->chunkCap = numChunks;
    vol->nextChunk = 0;
    vol->numLightingRequests = 0;
    vol->lightingRequestCap = numChunks;
    vol->numVoxelNodes = 0;
    vol->gpuMap = g->dFlags;
    vol->gpuChunks = g->dChunks;
    vol->gpuVoxels = g->dVoxels;

    // defaults, voxel.c:259-278
    vol->camPos = DNvec3{0.0f, 0.0f, 0.0f};
    vol->camOrient = DNvec3{0.0f, 0.0f, 0.0f};
    vol->camFOV = 90.0f;
    vol->camViewMode = 0;
    vol->sunDir = DNvec3{1.0f, 1.0f, 1.0f};
    vol->sunStrength = DNvec3{0.6f, 0.6f, 0.6f};
    vol->ambientLightStrength = DNvec3{0.01f, 0.01f, 0.01f};
    vol->diffuseBounceLimit = 5;
    vol->specBounceLimit = 2;
    vol->shadowSoftness = 10.0f;
    vol->useCubemap = false;
    vol->skyGradientBot = DNvec3{0.71f, 0.85f, 0.90f};
    vol->skyGradientTop = DNvec3{0.00f, 0.45f, 0.74f};
    vol->frameNum = 0;
    vol->lastTime = 123.456f;
    vol->numSpecularRays = 15;
    vol->numDirectRays = 0;
    return vol;
}

extern "C" void DN_delete_volume(DNvolume* vol)
{
    if(!vol) return;
    DNgpuVolume* g = vol->gpu;
    if(g)
    {
        if(g->stream) hipStreamSynchronize(g->stream);
        free_tile_buffers(g);
        dfree(g->dVoxels); dfree(g->dVoxelsOut); dfree(g->dMats); dfree(g->dNumRequests); dfree(g->dWorkCounter);
        dfree(g->dCounters); dfree(g->dPoints); dfree(g->dStage); dfree(g->dEdit); dfree(g->dCube); dfree(g->dDirect);
        if(g->hEdit) hipHostFree(g->hEdit);
        if(g->editEvents) hipEventDestroy(g->editEvents);
        for(int s = 0; s < 2; s++)
        {
            if(g->hStage[s]) hipHostFree(g->hStage[s]);
            if(g->stageDone[s]) hipEventDestroy(g->stageDone[s]);
        }
        if(g->hSmall) hipHostFree(g->hSmall);
        for(auto& t : g->freeTimers) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
        for(auto& t : g->pendingTimers) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
        if(g->stream) hipStreamDestroy(g->stream);
        delete g;
    }
    free(vol->map);
    free(vol->chunks);
    free(vol->materials);
    free(vol->lightingRequests);
    free(vol->gpuVoxelLayout);
    free(vol);
}

// ===========================================================================
// MEMORY (voxel.c:659-714)

extern "C" int DN_add_chunk(DNvolume* vol, DNivec3 pos)
{
    size_t i = vol->nextChunk;
    do
    {
        // the search reads one position per 4 KiB chunk, each on a page of its
        // own: fetch the positions a few slots ahead so the misses overlap
        const size_t ahead = (i + 8) % vol->chunkCap;
        __builtin_prefetch(&vol->chunks[ahead].pos);
        if(!DN_in_map_bounds(vol, vol->chunks[i].pos))
        {
            uint32_t mi = flat(vol, pos);
            vol->map[mi].chunkIndex = (uint32_t)i;
            vol->map[mi].flag = 1;
            vol->chunks[i].pos = pos;
            vol->nextChunk = (i == vol->chunkCap - 1) ? 0 : i + 1;
            mark_dirty(vol->gpu, mi);
            return (int)i;
        }
        i++;
        if(i >= vol->chunkCap) i = 0;
    } while(i != vol->nextChunk);

    size_t tiles = (size_t)vol->mapSize.x * vol->mapSize.y * vol->mapSize.z;
    size_t newCap = std::min(vol->chunkCap * 2, tiles);
    dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_NOTE, "automatically resizing chunk memory to accomodate %zu chunks (%zu bytes)",
           newCap, newCap * sizeof(DNchunk));
    i = vol->chunkCap;
    if(newCap <= vol->chunkCap || !DN_set_max_chunks(vol, newCap)) return -1;
    uint32_t mi = flat(vol, pos);
    vol->map[mi].chunkIndex = (uint32_t)i;
    vol->map[mi].flag = 1;
    vol->chunks[i].pos = pos;
    vol->nextChunk = (i == vol->chunkCap - 1) ? 0 : i + 1;
    mark_dirty(vol->gpu, mi);
    return (int)i;
}

extern "C" void DN_remove_chunk(DNvolume* vol, DNivec3 pos)
{
    uint32_t mi = flat(vol, pos);
    vol->map[mi].flag = 0;
    vol->nextChunk = vol->map[mi].chunkIndex;
    clear_chunk(vol, vol->map[mi].chunkIndex);
    mark_dirty(vol->gpu, mi);
}

// ===========================================================================
// SYNC (voxel.c:719-786): streaming state machine on the host mirror, one
// staged upload, then the device-side visible-chunk compaction.

namespace {

// Streaming half of DN_sync_gpu (voxel.c:1491-1640). Plans the state changes
// of the dirty tiles (removed chunks dropped, added chunks marked unloaded,
// updated chunks unloaded and re-requested, requested chunks loaded), stages
// the upload, and only then commits: if the staging buffer cannot be had,
// nothing changes and the next sync retries. A chunk to load takes a free pool
// node, else the node of the least recently used chunk (evicted: its tile goes
// back to "unloaded", and a ray that enters it requests it again), else the
// pool grows (doubling, up to chunkCap * 512 voxels and the pool limit), else
// the load waits for a later request. A device step that fails after the
// commit schedules a full device reset for the next sync.
void stream_chunks(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    if(g->resetPending)
    {
        if(!HIP_OK(hipMemsetAsync(g->dFlags, 0, sizeof(uint32_t) * g->tiles, g->stream), "map buffer reset")) return;
        g->resetPending = false;
        g->pool.reset(g->pool.cap);
        for(uint32_t t = 0; t < g->tiles; t++)
        {
            g->nodeClass[t] = -1;
            g->blocks[t] = 0;
            set_state(g, t, 0);
        }
        g->loadedBlocks = 0;
        g->layoutDirty = true;
        g->occDirty = true;
        g->requestAll = true;
        g->fullScan = true;
    }
    collect_ray_requests(vol);
    if(g->fullScan || g->requestAll)
    {
        for(uint32_t t = 0; t < g->tiles; t++)
            if(g->fullScan || (g->state[t] != 2 && vol->map[t].flag != 0)) mark_dirty(g, t);
    }
    std::sort(g->dirty.begin(), g->dirty.end());

    // plan: each dirty tile's next state (voxel.c:1491-1536)
    struct Plan { uint32_t t; uint8_t before, after; bool freeNode; };
    std::vector<Plan> plan;
    std::vector<uint32_t> loads;
    plan.reserve(g->dirty.size());
    for(uint32_t t : g->dirty)
    {
        const bool cpu = vol->map[t].flag != 0;
        const DNchunk* ch = cpu ? &vol->chunks[vol->map[t].chunkIndex] : nullptr;
        uint8_t st = g->state[t];
        Plan pl = {t, st, st, false};
        if(cpu && st == 0) st = 1;
        else if(!cpu && st != 0)
        {
            pl.freeNode = st == 2;
            st = 0;
        }
        if(st == 2 && ch && ch->updated) { pl.freeNode = true; st = 3; }
        if(g->requestAll && cpu && st == 1) st = 3;
        pl.after = st;
        if(st == 3 && cpu) loads.push_back(t);
        plan.push_back(pl);
    }

    // convert the chunks to load (_DN_chunk_to_gpu, on host threads)
    std::vector<DNchunkRec> recs(loads.size());
    std::vector<uint4> voxBuf(loads.size() * 512);
    std::vector<uint32_t> counts(loads.size());
    if(!loads.empty())
    {
        const DNmaterial* mats = vol->materials;
        unsigned nth = std::min<unsigned>(8, std::max<unsigned>(1, std::thread::hardware_concurrency()));
        if(loads.size() < 64) nth = 1;
        auto work = [&](unsigned w) {
            for(size_t i = w; i < loads.size(); i += nth)
            {
                const DNchunk& ch = vol->chunks[vol->map[loads[i]].chunkIndex];
                counts[i] = chunk_to_gpu(mats, ch, recs[i], &voxBuf[i * 512]);
            }
        };
        if(nth == 1) work(0);
        else
        {
            std::vector<std::thread> th;
            for(unsigned w = 0; w < nth; w++) th.emplace_back(work, w);
            for(auto& x : th) x.join();
        }
    }
    size_t totalVox = 0;
    for(uint32_t n : counts) totalVox += n;
    // at most one operation per planned tile plus one eviction per load
    const size_t maxOps = plan.size() + loads.size();
    if(maxOps == 0 && !g->occDirty)
    {
        g->requestAll = false;
        g->fullScan = false;
        return;
    }
    const size_t oRec = (sizeof(DNuploadOp) * maxOps + 127) & ~(size_t)127;
    const size_t oVox = oRec + sizeof(DNchunkRec) * recs.size();
    const size_t oOcc = oVox + sizeof(uint4) * totalVox;
    const size_t bytes = oOcc + 2 * sizeof(uint32_t) * g->occWords;
    const int slot = acquire_stage_slot(g);
    if(!ensure_stage(g, slot, bytes) || !g->dStage)
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "chunk streaming postponed: no staging memory for %zu bytes",
               bytes);
        return;                                  // nothing committed; the dirty tiles stay dirty
    }

    // commit the plan
    std::vector<DNuploadOp> ops;
    ops.reserve(maxOps);
    for(const Plan& pl : plan)
    {
        g->dirtyMark[pl.t] = 0;
        if(pl.freeNode) free_node(g, pl.t);
        set_state(g, pl.t, pl.after);
        if(vol->map[pl.t].flag != 0) vol->chunks[vol->map[pl.t].chunkIndex].updated = false;
        if(pl.after != pl.before && (pl.after == 0 || pl.after == 1)) ops.push_back(DNuploadOp{pl.t, pl.after, 0, 0});
    }
    g->dirty.clear();
    g->requestAll = false;
    g->fullScan = false;

    // pool nodes: free node, else evict the least recently used chunk, else grow, else wait.
    // Chunks loaded by this sync are marked state 3 until their node is found,
    // then 2, and are never eviction candidates (their device age is stale).
    EvictionQueue evq;
    for(uint32_t t : loads) g->dirtyMark[t] = 2;
    size_t voxOff = 0;
    for(size_t i = 0; i < loads.size(); i++)
    {
        const uint32_t t = loads[i], n = counts[i];
        vol->chunks[vol->map[t].chunkIndex].numVoxelsGpu = n;
        const int k = PoolAlloc::cls(n < 16 ? 16 : n);
        uint32_t start = g->pool.alloc(k);
        if(start == UINT32_MAX)
        {
            if(!evq.ready) build_eviction_queue(vol, evq);
            uint32_t e = pop_eviction(g, evq, k);
            if(e != UINT32_MAX)
            {
                free_node(g, e);
                set_state(g, e, 1);
                ops.push_back(DNuploadOp{e, 1, 0, 0});
                g->evictions++;
                start = g->pool.alloc(k);
            }
        }
        if(start == UINT32_MAX)
        {
            uint64_t want = std::max<uint64_t>((uint64_t)g->pool.cap * 2, (uint64_t)g->pool.cap + 512);
            want = std::min<uint64_t>(want, (uint64_t)vol->chunkCap * DN_CHUNK_LENGTH + 512);
            if(g->poolLimit) want = std::min<uint64_t>(want, g->poolLimit);
            want = std::min<uint64_t>(want, 0xFFFFFE00ull);
            if(want > g->pool.cap)
            {
                dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_NOTE,
                       "automatically resizing voxel buffer to accomodate %llu GPU voxels (%llu bytes)",
                       (unsigned long long)want, (unsigned long long)want * 16ull);
                if(grow_pool(vol, (uint32_t)want)) start = g->pool.alloc(k);
            }
        }
        if(start == UINT32_MAX)
        {
            set_state(g, t, 1);                  // waits for a later request (voxel.c:1607-1611)
            ops.push_back(DNuploadOp{t, 1, 0, 0});
            g->deferrals++;
            continue;
        }
        g->nodeStart[t] = start;
        g->nodeClass[t] = (int8_t)k;
        g->layoutDirty = true;
        recs[i].voxelIndex = start;
        g->blocks[t] = (n + 31u) / 32u;
        g->loadedBlocks += g->blocks[t];
        ops.push_back(DNuploadOp{t, 2, (uint32_t)i, (uint32_t)voxOff});
        set_state(g, t, 2);
        voxOff += n;
    }

    for(uint32_t t : loads) g->dirtyMark[t] = 0;

    // stage [ops][records][voxels][occLoaded][occPresent] and upload
    uint8_t* h = g->hStage[slot];
    memcpy(h, ops.data(), sizeof(DNuploadOp) * ops.size());
    if(!recs.empty()) memcpy(h + oRec, recs.data(), sizeof(DNchunkRec) * recs.size());
    uint4* hv = (uint4*)(h + oVox);
    for(const DNuploadOp& o : ops)
        if(o.mode == 2) memcpy(hv + o.voxOffset, &voxBuf[(size_t)o.recSlot * 512], sizeof(uint4) * counts[o.recSlot]);
    const size_t stagedVox = oVox + sizeof(uint4) * voxOff;
    memcpy(h + stagedVox, g->occLoaded.data(), sizeof(uint32_t) * g->occWords);
    memcpy(h + stagedVox + sizeof(uint32_t) * g->occWords, g->occPresent.data(), sizeof(uint32_t) * g->occWords);
    const size_t used = stagedVox + 2 * sizeof(uint32_t) * g->occWords;
    Timer tm = timer_begin(g, TK_UPLOAD);
    bool ok = HIP_OK(hipMemcpyAsync(g->dStage, h, used, hipMemcpyHostToDevice, g->stream), "chunk upload") &&
              HIP_OK(hipMemcpyAsync(g->dOccLoaded, g->dStage + stagedVox, sizeof(uint32_t) * g->occWords,
                                    hipMemcpyDeviceToDevice, g->stream), "occupancy upload") &&
              HIP_OK(hipMemcpyAsync(g->dOccPresent, g->dStage + stagedVox + sizeof(uint32_t) * g->occWords,
                                    sizeof(uint32_t) * g->occWords, hipMemcpyDeviceToDevice, g->stream),
                     "occupancy upload");
    if(ok && !ops.empty())
    {
        int e = dn_launch_upload(g->dFlags, g->dLastUsed, g->dChunks, g->dVoxels, g->dFresh,
                                 (const DNuploadOp*)g->dStage, (uint32_t)ops.size(), (const DNchunkRec*)(g->dStage + oRec),
                                 (const uint4*)(g->dStage + oVox), g->stream);
        if(e)
        {
            dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "chunk upload launch failed (%d)", e);
            ok = false;
        }
    }
    timer_end(g, tm);
    release_stage_slot(g, slot);
    g->upBytes += used;
    g->upChunks += recs.size();
    g->upVoxels += voxOff;
    if(!ok)
    {
        schedule_device_reset(g);
        return;
    }
    g->occDirty = false;
    rebuild_layout(vol);
}

}  // namespace

extern "C" void DN_sync_gpu(DNvolume* vol, DNmemOp op, int lightingSplit)
{
    DNgpuVolume* g = vol->gpu;
    if(lightingSplit < 1)
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "lightingSplit must be >= 1");
        lightingSplit = 1;
    }
    vol->frameNum++;
    if(vol->frameNum >= (uint32_t)lightingSplit) vol->frameNum = 0;

    if(op != DN_READ) stream_chunks(vol);

    DNvisibleParams vp;
    vp.flags = g->dFlags;
    vp.lastUsed = g->dLastUsed;
    vp.chunks = g->dChunks;
    vp.fresh = g->dFresh;
    vp.tiles = g->tiles;
    vp.split = op == DN_WRITE ? 0u : (uint32_t)lightingSplit;
    vp.frameNum = vol->frameNum;
    vp.blockSums = g->dBlockSums;
    vp.requests = g->dRequests;
    vp.numRequests = g->dNumRequests;
    Timer tm = timer_begin(g, TK_VISIBLE);
    int e = dn_launch_visible(vp, g->stream);
    timer_end(g, tm);
    if(e) dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "visible pass launch failed (%d)", e);
}

// ===========================================================================
// DRAWING (voxel.c:788-881)

namespace {
// QuickMath semantics (quickmath.h:199-203, 919-928, 1214-1333, 1037-1124),
// restated in plain float code.
float deg2rad(float d) { return d * 0.01745329251f; }
DNmat4 mat_identity()
{
    DNmat4 r;
    memset(&r, 0, sizeof(r));
    for(int i = 0; i < 4; i++) r.m[i][i] = 1.0f;
    return r;
}
DNmat4 mat_mult(const DNmat4& a, const DNmat4& b)   // a * b, column-major
{
    DNmat4 r;
    for(int c = 0; c < 4; c++)
    for(int row = 0; row < 4; row++)
    {
        float acc = a.m[0][row] * b.m[c][0];
        acc = acc + a.m[1][row] * b.m[c][1];
        acc = acc + a.m[2][row] * b.m[c][2];
        acc = acc + a.m[3][row] * b.m[c][3];
        r.m[c][row] = acc;
    }
    return r;
}
DNmat4 mat_inv(const DNmat4& mat)   // cofactor expansion, as quickmath.h:1037-1124
{
    float a = mat.m[0][0], b = mat.m[0][1], c = mat.m[0][2], d = mat.m[0][3], e = mat.m[1][0], f = mat.m[1][1],
          g = mat.m[1][2], h = mat.m[1][3], i = mat.m[2][0], j = mat.m[2][1], k = mat.m[2][2], l = mat.m[2][3],
          m = mat.m[3][0], n = mat.m[3][1], o = mat.m[3][2], p = mat.m[3][3];
    DNmat4 r;
    float t0 = k * p - o * l, t1 = j * p - n * l, t2 = j * o - n * k, t3 = i * p - m * l, t4 = i * o - m * k,
          t5 = i * n - m * j;
    r.m[0][0] = f * t0 - g * t1 + h * t2;
    r.m[1][0] = -(e * t0 - g * t3 + h * t4);
    r.m[2][0] = e * t1 - f * t3 + h * t5;
    r.m[3][0] = -(e * t2 - f * t4 + g * t5);
    r.m[0][1] = -(b * t0 - c * t1 + d * t2);
    r.m[1][1] = a * t0 - c * t3 + d * t4;
    r.m[2][1] = -(a * t1 - b * t3 + d * t5);
    r.m[3][1] = a * t2 - b * t4 + c * t5;
    t0 = g * p - o * h; t1 = f * p - n * h; t2 = f * o - n * g; t3 = e * p - m * h; t4 = e * o - m * g; t5 = e * n - m * f;
    r.m[0][2] = b * t0 - c * t1 + d * t2;
    r.m[1][2] = -(a * t0 - c * t3 + d * t4);
    r.m[2][2] = a * t1 - b * t3 + d * t5;
    r.m[3][2] = -(a * t2 - b * t4 + c * t5);
    t0 = g * l - k * h; t1 = f * l - j * h; t2 = f * k - j * g; t3 = e * l - i * h; t4 = e * k - i * g; t5 = e * j - i * f;
    r.m[0][3] = -(b * t0 - c * t1 + d * t2);
    r.m[1][3] = a * t0 - c * t3 + d * t4;
    r.m[2][3] = -(a * t1 - b * t3 + d * t5);
    r.m[3][3] = a * t2 - b * t4 + c * t5;
    float det = 1.0f / (a * r.m[0][0] + b * r.m[1][0] + c * r.m[2][0] + d * r.m[3][0]);
    for(int x = 0; x < 4; x++)
        for(int y = 0; y < 4; y++) r.m[x][y] = r.m[x][y] * det;
    return r;
}
DNvec3 v3(float x, float y, float z) { return DNvec3{x, y, z}; }
DNvec3 v3norm(DNvec3 v)
{
    float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if(len == 0.0f) return v3(0, 0, 0);
    float inv = 1.0f / len;
    return v3(v.x * inv, v.y * inv, v.z * inv);
}
DNvec3 v3cross(DNvec3 a, DNvec3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
void rotate_euler3(DNvec3 angles, float r[3][3])
{
    float sx = sinf(deg2rad(angles.x)), cx = cosf(deg2rad(angles.x));
    float sy = sinf(deg2rad(angles.y)), cy = cosf(deg2rad(angles.y));
    float sz = sinf(deg2rad(angles.z)), cz = cosf(deg2rad(angles.z));
    r[0][0] = cy * cz; r[0][1] = cy * sz; r[0][2] = -sy;
    r[1][0] = sx * sy * cz - cx * sz; r[1][1] = sx * sy * sz + cx * cz; r[1][2] = sx * cy;
    r[2][0] = cx * sy * cz + sx * sz; r[2][1] = cx * sy * sz - sx * cz; r[2][2] = cx * cy;
}
DNvec3 mat3_mul(const float m[3][3], DNvec3 v)
{
    return v3(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
              m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
}
}  // namespace

extern "C" void DN_set_view_projection_matrices(DNvolume* vol, float aspectRatio, float nearPlane, float farPlane,
                                                DNmat4* view, DNmat4* projection)
{
    float rot[3][3];
    rotate_euler3(vol->camOrient, rot);
    DNvec3 front;
    float t = tanf(deg2rad(vol->camFOV * 0.5f));
    if(aspectRatio < 1.0f) front = mat3_mul(rot, v3(0.0f, 0.0f, aspectRatio / t));
    else front = mat3_mul(rot, v3(0.0f, 0.0f, 1.0f / t));
    // lookat(pos, pos + front, up), quickmath.h:1290-1333
    DNvec3 pos = vol->camPos;
    DNvec3 target = v3(pos.x + front.x, pos.y + front.y, pos.z + front.z);
    DNvec3 dir = v3norm(v3(pos.x - target.x, pos.y - target.y, pos.z - target.z));
    DNvec3 r = v3norm(v3cross(v3(0.0f, 1.0f, 0.0f), dir));
    DNvec3 u = v3cross(dir, r);
    DNmat4 rud = mat_identity();
    rud.m[0][0] = r.x; rud.m[1][0] = r.y; rud.m[2][0] = r.z;
    rud.m[0][1] = u.x; rud.m[1][1] = u.y; rud.m[2][1] = u.z;
    rud.m[0][2] = dir.x; rud.m[1][2] = dir.y; rud.m[2][2] = dir.z;
    DNmat4 tr = mat_identity();
    tr.m[3][0] = -pos.x; tr.m[3][1] = -pos.y; tr.m[3][2] = -pos.z;
    *view = mat_mult(rud, tr);
    // perspective(fov, 1/aspect, near, far), quickmath.h:1264-1282
    float aspect = 1.0f / aspectRatio;
    DNmat4 p;
    memset(&p, 0, sizeof(p));
    float scale = tanf(deg2rad(vol->camFOV * 0.5f)) * nearPlane;
    float right = aspect * scale, top = scale;
    p.m[0][0] = nearPlane / right;
    p.m[1][1] = nearPlane / top;
    p.m[2][2] = -(farPlane + nearPlane) / (farPlane - nearPlane);
    p.m[3][2] = -2.0f * farPlane * nearPlane / (farPlane - nearPlane);
    p.m[2][3] = -1.0f;
    *projection = p;
}

extern "C" void DNX_draw_matrices(DNmat4 view, DNmat4 projection, DNmat4* invView, DNmat4* invViewProj)
{
    DNmat4 centered = view;
    centered.m[3][0] = 0.0f; centered.m[3][1] = 0.0f; centered.m[3][2] = 0.0f;
    *invView = mat_inv(view);
    *invViewProj = mat_mult(mat_inv(centered), mat_inv(projection));
}

extern "C" void DN_draw(DNvolume* vol, DNtexture out, DNmat4 view, DNmat4 projection, int rasterColorTexture,
                        int rasterDepthTexture)
{
    DNgpuVolume* g = vol->gpu;
    if(rasterColorTexture >= 0 || rasterDepthTexture >= 0)
        dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR, "raster compositing is not supported; drawing without it");
    if(vol->useCubemap && !g->dCube)
        dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR,
               "useCubemap is set but no cubemap was given (DNX_set_sky_cubemap); using the gradient");
    if(!out.data || out.width == 0 || out.height == 0)
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "invalid output texture");
        return;
    }
    if(!upload_materials(vol)) return;
    DNdrawParams p;
    p.scene = scene_view(vol);
    DNmat4 iv, ivp;
    DNX_draw_matrices(view, projection, &iv, &ivp);
    memcpy(p.invView, &iv.m[0][0], sizeof(p.invView));
    memcpy(p.invViewProj, &ivp.m[0][0], sizeof(p.invViewProj));
    p.width = out.width;
    p.height = out.height;
    p.viewMode = vol->camViewMode;
    p.out = (float4*)out.data;
    p.hitOut = ((size_t)out.width * out.height <= g->drawHitCap) ? g->drawHitOut : nullptr;
    if(g->drawHitOut && !p.hitOut)
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "draw hit output smaller than the image; not written");
    p.counters = g->dCounters + 0 * CNT_SLOTS;
    p.quadCounter = g->dWorkCounter + 2;
    Timer tm = timer_begin(g, TK_DRAW);
    int e = dn_launch_draw(p, g->stream);
    timer_end(g, tm);
    if(e) dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR, "draw launch failed (%d)", e);
}

// ===========================================================================
// LIGHTING (voxel.c:883-952)

// Everything DN_update_lighting does before its launches (voxel.c:883-945):
// validates the limits, uploads materials and ray tables, and fills the kernel
// parameters. Shared by the single-volume and the batched lighting pass.
bool prepare_lighting(DNvolume* vol, int numDiffuseSamples, int maxDiffuseSamples, float time, DNlightParams& p,
                      uint32_t& maxRequests)
{
    DNgpuVolume* g = vol->gpu;
    if(vol->frameNum == 0) vol->lastTime = time;
    if(numDiffuseSamples < 0 || numDiffuseSamples > DN_MAX_SAMPLES || maxDiffuseSamples < 0 ||
       vol->diffuseBounceLimit > DN_MAX_BOUNCES)
    {
        dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR,
               "lighting limits: numDiffuseSamples in [0, %d], diffuseBounceLimit <= %d", DN_MAX_SAMPLES, DN_MAX_BOUNCES);
        return false;
    }
    if(!upload_materials(vol) || !ensure_points(vol)) return false;
    if(g->anySpecular && !g->dVoxelsOut)
        if(!dalloc(&g->dVoxelsOut, (size_t)g->pool.cap, "voxel output buffer")) return false;
    memset(&p, 0, sizeof(p));
    p.tiles = g->tiles;
    p.copyBack = g->anySpecular ? 1u : 0u;
    p.directOut = nullptr;
    p.directIn = nullptr;
    p.scene = scene_view(vol);
    p.requests = g->dRequests;
    p.numRequests = g->dNumRequests;
    p.workCounter = g->dWorkCounter;
    p.litBits = g->dLitBits;
    p.visProp = g->dVisProp;
    p.voxelsOut = g->anySpecular ? g->dVoxelsOut : g->dVoxels;
    p.debugValues = g->dbgValues;
    p.debugIndex = g->dbgIndex;
    p.debugSlots = (uint32_t)std::min<size_t>(g->dbgCap * 32u, 0xFFFFFFFFu);
    p.debugFresh = g->freshOut;
    p.freshSlots = (uint32_t)std::min<size_t>(g->freshCap * 32u, 0xFFFFFFFFu);
    p.counters = g->dCounters + 1 * CNT_SLOTS;
    p.numDiffuseSamples = (uint32_t)numDiffuseSamples;
    p.maxDiffuseSamples = (uint32_t)maxDiffuseSamples;
    p.diffuseBounceLimit = vol->diffuseBounceLimit;
    p.specBounceLimit = vol->specBounceLimit;
    DNvec3 sd = v3norm(vol->sunDir);
    p.sunDir[0] = sd.x; p.sunDir[1] = sd.y; p.sunDir[2] = sd.z;
    p.ambient[0] = vol->ambientLightStrength.x; p.ambient[1] = vol->ambientLightStrength.y;
    p.ambient[2] = vol->ambientLightStrength.z;
    p.shadowSoftness = vol->shadowSoftness;
    p.camPos[0] = vol->camPos.x; p.camPos[1] = vol->camPos.y; p.camPos[2] = vol->camPos.z;
    // the ray tables hold DN_MAX_POINTS directions each
    p.numSpecRays = std::min<uint32_t>(vol->numSpecularRays, DN_MAX_POINTS);
    p.numDirectRays = std::min<uint32_t>(vol->numDirectRays, DN_MAX_POINTS);
    p.specPoints = g->dPoints;
    frame_random(vol->lastTime, (uint32_t)numDiffuseSamples, vol->diffuseBounceLimit, p.rnd);
    // voxel-independent shadow rays, in the shader's FP32 operations (voxelLighting.comp:72-78)
    {
        auto ray = [](float x, float y, float z, float* o) {
            o[0] = x; o[1] = y; o[2] = z;
            o[3] = 1.0f / x; o[4] = 1.0f / y; o[5] = 1.0f / z;
        };
        auto soft = [&](const float* q, float* o) {
            float vx = p.sunDir[0] * p.shadowSoftness + q[0];
            float vy = p.sunDir[1] * p.shadowSoftness + q[1];
            float vz = p.sunDir[2] * p.shadowSoftness + q[2];
            float inv = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
            ray(vx * inv + 0.0001f, vy * inv + 0.0001f, vz * inv + 0.0001f, o);
        };
        for(uint32_t k = 0; k < vol->numDirectRays && k < DN_MAX_POINTS; k++) soft(&g->directPts[3 * k], p.directRays[k]);
        for(int i = 0; i < numDiffuseSamples && i < DN_MAX_SAMPLES; i++) soft(p.rnd.shadowDir[i], p.shadowRays[i]);
        ray(p.sunDir[0] + 0.0001f, p.sunDir[1] + 0.0001f, p.sunDir[2] + 0.0001f, p.firstShadow);
    }
    maxRequests = (uint32_t)std::min<uint64_t>(g->loadedBlocks, (uint64_t)g->tiles * 16u);
    return true;
}

extern "C" void DN_update_lighting(DNvolume* vol, int numDiffuseSamples, int maxDiffuseSamples, float time)
{
    DNgpuVolume* g = vol->gpu;
    DNlightParams p;
    uint32_t maxReq = 0;
    if(!prepare_lighting(vol, numDiffuseSamples, maxDiffuseSamples, time, p, maxReq)) return;
    // Direct rays traced ahead in direction-major waves (k_direct); voxels whose
    // rays depend on ray order (transparent voxels) are traced by the pass.
    const bool ahead = p.numDirectRays > 0 && p.numDiffuseSamples > 0 && maxReq > 0 && ensure_direct(g, maxReq);
    Timer tm = timer_begin(g, TK_LIGHT);
    int e = 0;
    if(ahead)
    {
        p.directOut = g->dDirect;
        p.directSlots = (uint32_t)std::min<size_t>(g->directCap, 0xFFFFFFFFu);
        e = dn_launch_direct(p, maxReq, g->stream);
        p.directOut = nullptr;
        p.directIn = g->dDirect;
    }
    if(!e) e = dn_launch_lighting(p, maxReq, g->stream);
    timer_end(g, tm);
    Timer te = timer_begin(g, TK_EPILOGUE);
    int e2 = dn_launch_lighting_epilogue(p, g->tiles, g->dRequests, g->dNumRequests, g->anySpecular ? 1 : 0, g->stream);
    timer_end(g, te);
    if(e || e2) dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR, "lighting launch failed (%d, %d)", e, e2);
}

bool trace_direct_terms(DNvolume* vol, DNlightParams& p, uint32_t maxRequests)
{
    DNgpuVolume* g = vol->gpu;
    if(p.numDirectRays == 0 || maxRequests == 0 || !ensure_direct(g, maxRequests)) return false;
    p.directOut = g->dDirect;
    p.directSlots = (uint32_t)std::min<size_t>(g->directCap, 0xFFFFFFFFu);
    const int e = dn_launch_direct(p, maxRequests, g->stream);
    p.directOut = nullptr;
    // k_direct's request counter, which the lighting pass's epilogue clears
    // otherwise (this runs outside a lighting pass and outside graph capture)
    const bool ok = e == 0 && HIP_OK(hipMemsetAsync(g->dWorkCounter + 4, 0, sizeof(uint32_t), g->stream), "counter");
    if(ok) p.directIn = g->dDirect;
    else p.directSlots = 0;
    return ok;
}

// ===========================================================================
// MAP SETTINGS (voxel.c:957-1096)

extern "C" bool DN_set_map_size(DNvolume* vol, DNuvec3 size)
{
    if(size.x == 0 || size.y == 0 || size.z == 0 || (uint64_t)size.x * size.y * size.z >= (1ull << 28))
    {
        dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_ERROR, "invalid map size");
        return false;
    }
    uint32_t tiles = size.x * size.y * size.z;
    DNchunkHandle* nm = (DNchunkHandle*)calloc(tiles, sizeof(DNchunkHandle));
    if(!nm)
    {
        dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_ERROR, "failed to reallocate memory for map");
        return false;
    }
    for(uint32_t z = 0; z < size.z; z++)
    for(uint32_t y = 0; y < size.y; y++)
    for(uint32_t x = 0; x < size.x; x++)
    {
        DNivec3 p = {(int32_t)x, (int32_t)y, (int32_t)z};
        if(DN_in_map_bounds(vol, p)) nm[x + size.x * (y + z * size.y)] = vol->map[flat(vol, p)];
    }
    free(vol->map);
    vol->map = nm;
    vol->mapSize = size;
    for(size_t i = 0; i < vol->chunkCap; i++)
        if(!DN_in_map_bounds(vol, vol->chunks[i].pos)) clear_chunk(vol, i);
    DNgpuVolume* g = vol->gpu;
    hipStreamSynchronize(g->stream);
    free_tile_buffers(g);
    if(!alloc_tile_buffers(g, tiles)) return false;
    g->pool.reset(g->pool.cap);
    g->layoutDirty = true;
    g->fullScan = true;
    vol->gpuMap = g->dFlags;
    vol->gpuChunks = g->dChunks;
    for(uint32_t t = 0; t < tiles; t++)
        if(vol->map[t].flag) vol->chunks[vol->map[t].chunkIndex].updated = true;
    return true;
}

extern "C" bool DN_set_max_chunks(DNvolume* vol, size_t num)
{
    DNchunk* nc = (DNchunk*)realloc(vol->chunks, sizeof(DNchunk) * num);
    if(!nc)
    {
        dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_ERROR, "failed to reallocate memory for chunks");
        return false;
    }
    vol->chunks = nc;
    for(size_t i = vol->chunkCap; i < num; i++) clear_chunk(vol, i);
    vol->chunkCap = num;
    return true;
}

extern "C" bool DN_set_max_voxels_gpu(DNvolume* vol, size_t num)
{
    if(num < vol->gpu->pool.cap || num > 0xFFFFFE00u || (vol->gpu->poolLimit && num > vol->gpu->poolLimit))
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "the voxel buffer can only grow (requested %zu, have %u)", num,
               vol->gpu->pool.cap);
        return false;
    }
    return grow_pool(vol, (uint32_t)num);
}

extern "C" bool DNX_set_voxel_pool_limit(DNvolume* vol, size_t maxVoxels)
{
    DNgpuVolume* g = vol->gpu;
    if(maxVoxels == 0) { g->poolLimit = 0; return true; }
    size_t lim = (maxVoxels + 511u) & ~(size_t)511u;
    if(lim < g->pool.cap || lim > 0xFFFFFE00u)
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "voxel pool limit %zu is below the pool's %u voxels or too large",
               maxVoxels, g->pool.cap);
        return false;
    }
    g->poolLimit = (uint32_t)lim;
    return true;
}

extern "C" void DNX_get_residency_stats(DNvolume* vol, uint64_t* evictions, uint64_t* deferrals, bool reset)
{
    DNgpuVolume* g = vol->gpu;
    if(evictions) *evictions = g->evictions;
    if(deferrals) *deferrals = g->deferrals;
    if(reset) g->evictions = g->deferrals = 0;
}

extern "C" bool DN_set_max_lighting_requests(DNvolume* vol, size_t num)
{
    uint32_t* nr = (uint32_t*)realloc(vol->lightingRequests, sizeof(uint32_t) * (num ? num : 1));
    if(!nr)
    {
        dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_ERROR, "failed to reallocate memory for lighting requests");
        return false;
    }
    vol->lightingRequests = nr;
    vol->lightingRequestCap = num;
    return true;
}

// ===========================================================================
// MAP UTILITY (voxel.c:1101-1272)

extern "C" bool DN_in_map_bounds(DNvolume* vol, DNivec3 pos)
{
    return pos.x >= 0 && pos.y >= 0 && pos.z >= 0 && (uint32_t)pos.x < vol->mapSize.x &&
           (uint32_t)pos.y < vol->mapSize.y && (uint32_t)pos.z < vol->mapSize.z;
}

extern "C" bool DN_in_chunk_bounds(DNivec3 pos)
{
    return pos.x < DN_CHUNK_SIZE && pos.y < DN_CHUNK_SIZE && pos.z < DN_CHUNK_SIZE && pos.x >= 0 && pos.y >= 0 &&
           pos.z >= 0;
}

extern "C" DNcompressedVoxel DN_get_compressed_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos)
{
    return vol->chunks[vol->map[flat(vol, mapPos)].chunkIndex].voxels[chunkPos.x][chunkPos.y][chunkPos.z];
}

extern "C" DNvoxel DN_get_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos)
{
    return DN_decompress_voxel(DN_get_compressed_voxel(vol, mapPos, chunkPos));
}

extern "C" void DN_set_compressed_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos, DNcompressedVoxel voxel)
{
    uint32_t mi = flat(vol, mapPos);
    if(vol->map[mi].flag == 0)
    {
        if(material_of(voxel.normal) == DN_MATERIAL_EMPTY) return;
        if(DN_add_chunk(vol, mapPos) < 0) return;
    }
    uint32_t ci = vol->map[mi].chunkIndex;
    DNchunk& ch = vol->chunks[ci];
    uint32_t oldMat = material_of(ch.voxels[chunkPos.x][chunkPos.y][chunkPos.z].normal);
    uint32_t newMat = material_of(voxel.normal);
    if(oldMat == DN_MATERIAL_EMPTY && newMat != DN_MATERIAL_EMPTY) ch.numVoxels++;
    else if(oldMat != DN_MATERIAL_EMPTY && newMat == DN_MATERIAL_EMPTY)
    {
        ch.numVoxels--;
        if(ch.numVoxels == 0)
        {
            DN_remove_chunk(vol, mapPos);
            return;
        }
    }
    ch.voxels[chunkPos.x][chunkPos.y][chunkPos.z] = voxel;
    ch.updated = true;
    mark_dirty(vol->gpu, mi);
}

extern "C" void DN_set_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos, DNvoxel voxel)
{
    DN_set_compressed_voxel(vol, mapPos, chunkPos, DN_compress_voxel(voxel));
}

extern "C" bool DN_does_chunk_exist(DNvolume* vol, DNivec3 pos) { return vol->map[flat(vol, pos)].flag >= 1; }

extern "C" bool DN_does_voxel_exist(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos)
{
    return material_of(DN_get_compressed_voxel(vol, mapPos, chunkPos).normal) != DN_MATERIAL_EMPTY;
}

extern "C" void DN_remove_voxel(DNvolume* vol, DNivec3 mapPos, DNivec3 chunkPos)
{
    uint32_t mi = flat(vol, mapPos);
    uint32_t ci = vol->map[mi].chunkIndex;
    if(DN_does_voxel_exist(vol, mapPos, chunkPos))
    {
        vol->chunks[ci].numVoxels--;
        if(vol->chunks[ci].numVoxels == 0)
        {
            DN_remove_chunk(vol, mapPos);
            return;
        }
    }
    vol->chunks[ci].voxels[chunkPos.x][chunkPos.y][chunkPos.z].normal = UINT32_MAX;
    vol->chunks[ci].updated = true;
    mark_dirty(vol->gpu, mi);
}

// CPU pick ray (voxel.c:1195-1272): voxel-unit DDA with strict-less ties,
// z winning ties (unlike the GPU DDA's multi-axis steps).
extern "C" bool DN_step_map(DNvolume* vol, DNvec3 rayDir, DNvec3 rayPos, int maxSteps, DNivec3* hitPos,
                            DNvoxel* hitVoxel, DNivec3* hitNormal)
{
    auto sgn = [](float n) { return n > 0 ? 1 : (n < 0 ? -1 : 0); };
    *hitNormal = DNivec3{-1000, -1000, -1000};
    rayPos = v3(rayPos.x * DN_CHUNK_SIZE, rayPos.y * DN_CHUNK_SIZE, rayPos.z * DN_CHUNK_SIZE);
    DNvec3 inv = v3(1.0f / rayDir.x, 1.0f / rayDir.y, 1.0f / rayDir.z);
    int sx = sgn(rayDir.x), sy = sgn(rayDir.y), sz = sgn(rayDir.z);
    DNivec3 pos = {(int32_t)floor(rayPos.x), (int32_t)floor(rayPos.y), (int32_t)floor(rayPos.z)};
    DNvec3 dd = v3(fabsf(inv.x), fabsf(inv.y), fabsf(inv.z));
    DNvec3 side;
    side.x = (float)((sx * (pos.x - rayPos.x) + (sx * 0.5) + 0.5) * dd.x);
    side.y = (float)((sy * (pos.y - rayPos.y) + (sy * 0.5) + 0.5) * dd.y);
    side.z = (float)((sz * (pos.z - rayPos.z) + (sz * 0.5) + 0.5) * dd.z);
    for(int steps = 0; steps < maxSteps; steps++)
    {
        DNivec3 mp, cp;
        DN_separate_position(pos, &mp, &cp);
        if(pos.x < 0) mp.x--;
        if(pos.y < 0) mp.y--;
        if(pos.z < 0) mp.z--;
        if(DN_in_map_bounds(vol, mp) && DN_does_chunk_exist(vol, mp) && DN_in_chunk_bounds(cp) &&
           DN_does_voxel_exist(vol, mp, cp))
        {
            *hitVoxel = DN_get_voxel(vol, mp, cp);
            *hitPos = pos;
            return true;
        }
        if(side.x < side.y)
        {
            if(side.x < side.z) { side.x += dd.x; pos.x += sx; *hitNormal = DNivec3{-sx, 0, 0}; }
            else { side.z += dd.z; pos.z += sz; *hitNormal = DNivec3{0, 0, -sz}; }
        }
        else
        {
            if(side.y < side.z) { side.y += dd.y; pos.y += sy; *hitNormal = DNivec3{0, -sy, 0}; }
            else { side.z += dd.z; pos.z += sz; *hitNormal = DNivec3{0, 0, -sz}; }
        }
    }
    return false;
}

// ===========================================================================
// GENERAL UTILITY (voxel.c:1277-1316)

extern "C" void DN_separate_position(DNivec3 pos, DNivec3* mapPos, DNivec3* chunkPos)
{
    *mapPos = DNivec3{pos.x / DN_CHUNK_SIZE, pos.y / DN_CHUNK_SIZE, pos.z / DN_CHUNK_SIZE};
    *chunkPos = DNivec3{pos.x % DN_CHUNK_SIZE, pos.y % DN_CHUNK_SIZE, pos.z % DN_CHUNK_SIZE};
}

extern "C" DNvec3 DN_cam_dir(DNvec3 orient)
{
    float rot[3][3];
    rotate_euler3(orient, rot);
    return mat3_mul(rot, v3(0.0f, 0.0f, 1.0f));
}

extern "C" DNcompressedVoxel DN_compress_voxel(DNvoxel voxel)
{
    // min(x, 1) = x < 1 ? x : 1 and max(x, -1) = x > -1 ? x : -1, so a NaN
    // component (fromCenter / 0 at a sphere's centre voxel) quantises as 1.
    DNvec3 n = voxel.normal;
    auto clamp1 = [](float x) { x = x < 1.0f ? x : 1.0f; return x > -1.0f ? x : -1.0f; };
    n.x = clamp1(n.x);
    n.y = clamp1(n.y);
    n.z = clamp1(n.z);
    uint32_t nx = (uint32_t)(((int)(n.x * 255.0f) + 255) / 2), ny = (uint32_t)(((int)(n.y * 255.0f) + 255) / 2),
             nz = (uint32_t)(((int)(n.z * 255.0f) + 255) / 2);
    DNcompressedVoxel r;
    r.normal = ((uint32_t)voxel.material << 24) | (nx << 16) | (ny << 8) | nz;
    r.albedo = ((uint32_t)voxel.albedo.r << 24) | ((uint32_t)voxel.albedo.g << 16) | ((uint32_t)voxel.albedo.b << 8);
    return r;
}

extern "C" DNvoxel DN_decompress_voxel(DNcompressedVoxel voxel)
{
    const float inv255 = 1.0f / 255.0f;
    int nx = (int)((voxel.normal >> 16) & 0xFF) * 2 - 255, ny = (int)((voxel.normal >> 8) & 0xFF) * 2 - 255,
        nz = (int)(voxel.normal & 0xFF) * 2 - 255;
    DNvoxel r;
    r.normal = v3((float)nx * inv255, (float)ny * inv255, (float)nz * inv255);
    r.material = (uint8_t)(voxel.normal >> 24);
    r.albedo = DNcolor{(uint8_t)((voxel.albedo >> 24) & 0xFF), (uint8_t)((voxel.albedo >> 16) & 0xFF),
                       (uint8_t)((voxel.albedo >> 8) & 0xFF)};
    return r;
}

namespace {
// The tile box a sphere edit visits (voxelShapes.c:248-275: floor/ceil of the
// bounds, one more voxel each way when carving, DN_separate_position's
// truncating split, clamped to the map) and its FP32 radii.
struct SphereBox {
    DNivec3 lo, hi;
    float r2, r12;
    bool shell;                    // carve with flipped normals: the next shell is touched too
};

SphereBox sphere_box(const DNvolume* vol, uint8_t material, bool flipNormals, DNvec3 c, float r)
{
    DNvec3 mn = v3(c.x - r, c.y - r, c.z - r), mx = v3(c.x + r, c.y + r, c.z + r);
    DNivec3 iMin = {(int)floorf(mn.x), (int)floorf(mn.y), (int)floorf(mn.z)};
    DNivec3 iMax = {(int)ceilf(mx.x), (int)ceilf(mx.y), (int)ceilf(mx.z)};
    if(material == DN_MATERIAL_EMPTY)
    {
        iMin = DNivec3{iMin.x - 1, iMin.y - 1, iMin.z - 1};
        iMax = DNivec3{iMax.x + 1, iMax.y + 1, iMax.z + 1};
    }
    SphereBox b;
    DNivec3 cmin, cmax;
    DN_separate_position(iMin, &b.lo, &cmin);
    DN_separate_position(iMax, &b.hi, &cmax);
    b.lo.x = std::max(b.lo.x, 0); b.lo.y = std::max(b.lo.y, 0); b.lo.z = std::max(b.lo.z, 0);
    b.hi.x = std::min(b.hi.x, (int)vol->mapSize.x - 1);
    b.hi.y = std::min(b.hi.y, (int)vol->mapSize.y - 1);
    b.hi.z = std::min(b.hi.z, (int)vol->mapSize.z - 1);
    b.r2 = r * r;
    b.r12 = (r + 1.0f) * (r + 1.0f);
    b.shell = material == DN_MATERIAL_EMPTY && flipNormals;
    return b;
}

// A voxel is touched only when its FP32 squared distance is below r2 (or r12
// for the normal-flipping shell). Rounding is monotone, so the voxel of a tile
// (or of a row) nearest to the centre along each axis has the smallest FP32
// distance of all of them: tiles and rows whose nearest voxel fails both tests
// can be skipped without changing which voxels are written or in what order.
inline bool sphere_touches(const SphereBox& b, float d2) { return d2 < b.r2 || (b.shell && d2 < b.r12); }

inline float nearest_offset(int base, float cc)
{
    float m = fabsf((float)base - cc);
    for(int k = 1; k < 8; k++) m = std::min(m, fabsf((float)(base + k) - cc));
    return m;
}
}  // namespace

namespace {
// QuickMath's quaternion_to_mat4 (quickmath.h:1622-1650): rotation part of a
// unit quaternion, written into an identity matrix (column-major m[col][row]).
DNmat4 quaternion_mat(DNquaternion q)
{
    DNmat4 r = mat_identity();
    float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    float xx2 = q.x * x2, xy2 = q.x * y2, xz2 = q.x * z2;
    float yy2 = q.y * y2, yz2 = q.y * z2, zz2 = q.z * z2;
    float sx2 = q.w * x2, sy2 = q.w * y2, sz2 = q.w * z2;
    r.m[0][0] = 1.0f - (yy2 + zz2); r.m[0][1] = xy2 - sz2;          r.m[0][2] = xz2 + sy2;
    r.m[1][0] = xy2 + sz2;          r.m[1][1] = 1.0f - (xx2 + zz2); r.m[1][2] = yz2 - sx2;
    r.m[2][0] = xz2 - sy2;          r.m[2][1] = yz2 + sx2;          r.m[2][2] = 1.0f - (xx2 + yy2);
    return r;
}
}  // namespace

// voxelShapes.c:239-333. Visits the sphere's tile box in z, y, x order and each
// tile's voxels in z, y, x order; the tile's chunk-present flag is read once
// per tile, as the reference copies the map tile before its voxel loop. With
// `func`, the voxel placed (and a flipped rim voxel) is what func returns for
// the voxel's position, its computed normal, the voxel (the argument, or the
// old voxel on the rim), the sphere's bounding box and the inverse of
// translate(c) * rotation(orient) (voxelShapes.c:242-246, 297-303, 316-321).
extern "C" void DN_shape_sphere(DNvolume* vol, DNvoxel voxel, bool flipNormals, DNvec3 c, float r, DNquaternion orient,
                                VoxelTransformFunc func, void* userData)
{
    DNmat4 invTransform = mat_identity();
    if(func)
    {
        DNmat4 translate = mat_identity();
        translate.m[3][0] = c.x; translate.m[3][1] = c.y; translate.m[3][2] = c.z;
        invTransform = mat_inv(mat_mult(translate, quaternion_mat(orient)));
    }
    const DNvec3 shapeMin = v3(c.x - r, c.y - r, c.z - r), shapeMax = v3(c.x + r, c.y + r, c.z + r);
    const SphereBox b = sphere_box(vol, voxel.material, flipNormals, c, r);
    const DNivec3 mapMin = b.lo, mapMax = b.hi;
    const float r2 = b.r2, r12 = b.r12;
    auto touched = [&](float d2) { return sphere_touches(b, d2); };
    for(int mZ = mapMin.z; mZ <= mapMax.z; mZ++)
    for(int mY = mapMin.y; mY <= mapMax.y; mY++)
    for(int mX = mapMin.x; mX <= mapMax.x; mX++)
    {
        DNivec3 mp = {mX, mY, mZ};
        const float nx = nearest_offset(mX * 8, c.x), ny = nearest_offset(mY * 8, c.y), nz = nearest_offset(mZ * 8, c.z);
        if(!touched(nx * nx + ny * ny + nz * nz)) continue;
        const bool tileExists = vol->map[flat(vol, mp)].flag != 0;
        for(int cZ = 0; cZ < 8; cZ++)
        for(int cY = 0; cY < 8; cY++)
        {
            const float fy = (float)(mY * 8 + cY) - c.y, fz = (float)(mZ * 8 + cZ) - c.z;
            if(!touched(nx * nx + fy * fy + fz * fz)) continue;
        for(int cX = 0; cX < 8; cX++)
        {
            DNivec3 cp = {cX, cY, cZ};
            DNvec3 pos = v3((float)(mX * 8 + cX), (float)(mY * 8 + cY), (float)(mZ * 8 + cZ));
            DNvec3 fc = v3(pos.x - c.x, pos.y - c.y, pos.z - c.z);
            float d2 = fc.x * fc.x + fc.y * fc.y + fc.z * fc.z;
            if(d2 < r2)
            {
                if(tileExists && DN_does_voxel_exist(vol, mp, cp) && voxel.material != DN_MATERIAL_EMPTY) continue;
                DNvoxel fin = voxel;
                if(voxel.material != DN_MATERIAL_EMPTY)
                {
                    float mxn = fmaxf(fmaxf(fabsf(fc.x), fabsf(fc.y)), fabsf(fc.z));
                    float s = 1.0f / mxn;
                    DNvec3 n = v3(fc.x * s, fc.y * s, fc.z * s);
                    if(func) fin = func(pos, n, voxel, shapeMin, shapeMax, invTransform, userData);
                    else fin.normal = n;
                }
                DN_set_voxel(vol, mp, cp, fin);
            }
            else if(voxel.material == DN_MATERIAL_EMPTY && flipNormals && d2 < r12)
            {
                if(!tileExists || !DN_does_voxel_exist(vol, mp, cp)) continue;
                DNvoxel old = DN_get_voxel(vol, mp, cp);
                float mxn = fmaxf(fmaxf(fabsf(fc.x), fabsf(fc.y)), fabsf(fc.z));
                float s = -1.0f / mxn;
                DNvec3 n = v3(fc.x * s, fc.y * s, fc.z * s);
                if(func) old = func(pos, n, old, shapeMin, shapeMax, invTransform, userData);
                else old.normal = n;
                DN_set_voxel(vol, mp, cp, old);
            }
        }
        }
    }
}

// ===========================================================================
// FILE IO (voxel.c:296-654): the .voxvol volume file. Layout: map size (3 x
// u32), chunk capacity (u64), per chunk slot a u16 byte count and the
// compressed chunk, 256 raw 32-byte materials, then camera, lighting and sky
// parameters. A compressed chunk is its position (3 x i32), and for a slot in
// use a normal palette and an albedo palette (u8 count + 3-byte entries, in
// first-seen order, kept only while smaller than half the chunk's voxel count)
// followed by material runs over the voxels in x-fastest order: material, run
// length (<= 255), and per solid voxel its normal and albedo, each a palette
// index or 3 raw bytes.

namespace {
struct Rgb { uint8_t v[3]; };
inline bool same_rgb(const Rgb& a, const Rgb& b) { return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2]; }
inline Rgb normal_rgb(uint32_t w) { return Rgb{{(uint8_t)(w >> 16), (uint8_t)(w >> 8), (uint8_t)w}}; }
inline Rgb albedo_rgb(uint32_t w) { return Rgb{{(uint8_t)(w >> 24), (uint8_t)(w >> 16), (uint8_t)(w >> 8)}}; }

int find_rgb(const Rgb* pal, int n, const Rgb& c)
{
    for(int i = 0; i < n; i++)
        if(same_rgb(pal[i], c)) return i;
    return -1;
}

// _DN_compress_chunk (voxel.c:300-430): appends the chunk's record to `out`.
void encode_chunk(const DNvolume* vol, const DNchunk& ch, std::vector<uint8_t>& out)
{
    auto put = [&](const void* p, size_t n) {
        const uint8_t* b = (const uint8_t*)p;
        out.insert(out.end(), b, b + n);
    };
    put(&ch.pos, sizeof(DNivec3));
    if(!DN_in_map_bounds(const_cast<DNvolume*>(vol), ch.pos)) return;
    const int half = (int)(ch.numVoxels / 2);
    Rgb npal[DN_CHUNK_LENGTH / 2], apal[DN_CHUNK_LENGTH / 2];
    int nn = 0, na = 0;
    for(int i = 0; i < DN_CHUNK_LENGTH; i++)
    {
        const DNcompressedVoxel& v = ch.voxels[i & 7][(i >> 3) & 7][i >> 6];
        if(material_of(v.normal) == DN_MATERIAL_EMPTY) continue;
        const Rgb n = normal_rgb(v.normal), a = albedo_rgb(v.albedo);
        if(nn < half && find_rgb(npal, nn, n) < 0) npal[nn++] = n;
        if(na < half && find_rgb(apal, na, a) < 0) apal[na++] = a;
    }
    if(!(nn < half)) nn = 0;
    if(!(na < half)) na = 0;
    out.push_back((uint8_t)nn);
    put(npal, 3u * nn);
    out.push_back((uint8_t)na);
    put(apal, 3u * na);
    for(int i = 0; i < DN_CHUNK_LENGTH;)
    {
        const uint32_t mat = material_of(ch.voxels[i & 7][(i >> 3) & 7][i >> 6].normal);
        out.push_back((uint8_t)mat);
        const size_t at = out.size();
        out.push_back(0);
        int num = 0, j = i;
        for(; j < DN_CHUNK_LENGTH && num < 255; j++)
        {
            const DNcompressedVoxel& v = ch.voxels[j & 7][(j >> 3) & 7][j >> 6];
            if(material_of(v.normal) != mat) break;
            num++;
            if(mat == DN_MATERIAL_EMPTY) continue;
            const Rgb n = normal_rgb(v.normal), a = albedo_rgb(v.albedo);
            if(nn) out.push_back((uint8_t)find_rgb(npal, nn, n)); else put(n.v, 3);
            if(na) out.push_back((uint8_t)find_rgb(apal, na, a)); else put(a.v, 3);
        }
        out[at] = (uint8_t)num;
        i = j;
    }
}

// Bounds-checked little reader over a file image.
struct Reader {
    const uint8_t* p;
    size_t n, at = 0;
    bool ok = true;
    bool take(void* dst, size_t k)
    {
        if(!ok || k > n - at) { ok = false; return false; }
        memcpy(dst, p + at, k);
        at += k;
        return true;
    }
    uint8_t byte() { uint8_t b = 0; take(&b, 1); return b; }
};

// _DN_decompress_chunk (voxel.c:433-518) into a cleared slot. Returns false on
// a malformed record (the reference would read past it).
bool decode_chunk(const DNvolume* vol, const uint8_t* mem, size_t size, DNchunk& ch)
{
    Reader r{mem, size};
    r.take(&ch.pos, sizeof(DNivec3));
    ch.updated = false;
    ch.numVoxels = 0;
    ch.numVoxelsGpu = 0;
    if(!r.ok) return false;
    if(!DN_in_map_bounds(const_cast<DNvolume*>(vol), ch.pos)) return r.at == size;
    Rgb npal[256], apal[256];
    const int nn = r.byte();
    r.take(npal, 3u * nn);
    const int na = r.byte();
    r.take(apal, 3u * na);
    int done = 0;
    while(r.ok && done < DN_CHUNK_LENGTH)
    {
        const uint32_t mat = r.byte();
        const int num = r.byte();
        if(done + num > DN_CHUNK_LENGTH) return false;
        for(int i = done; i < done + num && r.ok; i++)
        {
            DNcompressedVoxel& v = ch.voxels[i & 7][(i >> 3) & 7][i >> 6];
            if(mat == DN_MATERIAL_EMPTY)
            {
                v.normal = UINT32_MAX;
                continue;
            }
            Rgb n, a;
            if(nn) { const int k = r.byte(); if(k >= nn) return false; n = npal[k]; } else r.take(n.v, 3);
            if(na) { const int k = r.byte(); if(k >= na) return false; a = apal[k]; } else r.take(a.v, 3);
            v.normal = (mat << 24) | ((uint32_t)n.v[0] << 16) | ((uint32_t)n.v[1] << 8) | n.v[2];
            v.albedo = ((uint32_t)a.v[0] << 24) | ((uint32_t)a.v[1] << 16) | ((uint32_t)a.v[2] << 8);
            ch.numVoxels++;
        }
        done += num;
    }
    return r.ok && r.at == size;
}

// The parameter block after the materials (voxel.c:577-591), in file order.
template <typename F> void volume_params(DNvolume* vol, F&& io)
{
    io(&vol->camPos, sizeof(DNvec3)); io(&vol->camOrient, sizeof(DNvec3));
    io(&vol->camFOV, sizeof(float)); io(&vol->camViewMode, sizeof(uint32_t));
    io(&vol->sunDir, sizeof(DNvec3)); io(&vol->sunStrength, sizeof(DNvec3));
    io(&vol->ambientLightStrength, sizeof(DNvec3)); io(&vol->diffuseBounceLimit, sizeof(uint32_t));
    io(&vol->specBounceLimit, sizeof(uint32_t)); io(&vol->shadowSoftness, sizeof(float));
    io(&vol->skyGradientBot, sizeof(DNvec3)); io(&vol->skyGradientTop, sizeof(DNvec3));
}
}  // namespace

// voxel.c:520-593. The file is read whole and every field is bounds-checked: a
// truncated or malformed file gives an error message and NULL.
extern "C" DNvolume* DN_load_volume(const char* filePath, unsigned int minChunks)
{
    FILE* f = filePath ? fopen(filePath, "rb") : nullptr;
    if(!f)
    {
        dn_msg(DN_MESSAGE_FILE_IO, DN_MESSAGE_ERROR, "failed to open file \"%s\" for reading", filePath ? filePath : "");
        return nullptr;
    }
    std::vector<uint8_t> data;
    uint8_t buf[1 << 16];
    for(size_t k; (k = fread(buf, 1, sizeof(buf), f)) > 0;) data.insert(data.end(), buf, buf + k);
    fclose(f);
    Reader r{data.data(), data.size()};
    DNuvec3 mapSize;
    uint64_t chunkCap = 0;
    r.take(&mapSize, sizeof(DNuvec3));
    r.take(&chunkCap, sizeof(uint64_t));
    // every slot takes at least its 2-byte count: bounds the allocation
    if(!r.ok || chunkCap == 0 || chunkCap > (data.size() - r.at) / 2)
    {
        dn_msg(DN_MESSAGE_FILE_IO, DN_MESSAGE_ERROR, "\"%s\" is not a volume file (bad header)", filePath);
        return nullptr;
    }
    DNvolume* vol = DN_create_volume(mapSize, minChunks);
    if(!vol) return nullptr;
    if(!DN_set_max_chunks(vol, (size_t)chunkCap)) { DN_delete_volume(vol); return nullptr; }
    if(vol->nextChunk >= vol->chunkCap) vol->nextChunk = 0;
    for(size_t i = 0; i < chunkCap; i++)
    {
        uint16_t size = 0;
        if(!r.take(&size, sizeof(uint16_t)) || size > data.size() - r.at ||
           !decode_chunk(vol, data.data() + r.at, size, vol->chunks[i]))
        {
            dn_msg(DN_MESSAGE_FILE_IO, DN_MESSAGE_ERROR, "\"%s\": malformed chunk record %zu", filePath, i);
            DN_delete_volume(vol);
            return nullptr;
        }
        r.at += size;
        if(DN_in_map_bounds(vol, vol->chunks[i].pos))
        {
            const uint32_t mi = flat(vol, vol->chunks[i].pos);
            vol->map[mi].flag = 1;
            vol->map[mi].chunkIndex = (uint32_t)i;
            mark_dirty(vol->gpu, mi);
        }
    }
    r.take(vol->materials, sizeof(DNmaterial) * DN_MAX_MATERIALS);
    volume_params(vol, [&](void* p, size_t k) { r.take(p, k); });
    if(!r.ok)
    {
        dn_msg(DN_MESSAGE_FILE_IO, DN_MESSAGE_ERROR, "\"%s\" is truncated", filePath);
        DN_delete_volume(vol);
        return nullptr;
    }
    return vol;
}

// voxel.c:595-654. The file image is built in memory and written at once.
extern "C" bool DN_save_volume(const char* filePath, DNvolume* vol)
{
    std::vector<uint8_t> out, rec;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* b = (const uint8_t*)p;
        out.insert(out.end(), b, b + n);
    };
    const uint64_t cap = vol->chunkCap;
    put(&vol->mapSize, sizeof(DNuvec3));
    put(&cap, sizeof(uint64_t));
    for(size_t i = 0; i < vol->chunkCap; i++)
    {
        rec.clear();
        encode_chunk(vol, vol->chunks[i], rec);
        const uint16_t size = (uint16_t)rec.size();
        put(&size, sizeof(uint16_t));
        put(rec.data(), rec.size());
    }
    put(vol->materials, sizeof(DNmaterial) * DN_MAX_MATERIALS);
    volume_params(vol, [&](void* p, size_t k) { put(p, k); });
    FILE* f = filePath ? fopen(filePath, "wb") : nullptr;
    if(!f)
    {
        dn_msg(DN_MESSAGE_FILE_IO, DN_MESSAGE_ERROR, "failed to open file \"%s\" for writing", filePath ? filePath : "");
        return false;
    }
    const bool ok = fwrite(out.data(), 1, out.size(), f) == out.size();
    if(fclose(f) != 0 || !ok)
    {
        dn_msg(DN_MESSAGE_FILE_IO, DN_MESSAGE_ERROR, "failed to write \"%s\"", filePath);
        return false;
    }
    return true;
}

namespace {
inline size_t align16(size_t x) { return (x + 15u) & ~(size_t)15u; }

bool ensure_edit_stage(DNgpuVolume* g, size_t hostBytes, size_t devBytes)
{
    if(g->hEditCap < hostBytes)
    {
        if(g->hEdit) hipHostFree(g->hEdit);
        g->hEdit = nullptr;
        g->hEditCap = 0;
        size_t cap = std::max(hostBytes + hostBytes / 2, (size_t)1 << 20);
        if(!HIP_OK(hipHostMalloc((void**)&g->hEdit, cap, hipHostMallocDefault), "pinned edit staging")) return false;
        g->hEditCap = cap;
    }
    if(g->dEditCap < devBytes)
    {
        dfree(g->dEdit);
        g->dEditCap = 0;
        size_t cap = std::max(devBytes + devBytes / 2, (size_t)1 << 20);
        if(!dalloc(&g->dEdit, cap, "device edit staging")) return false;
        g->dEditCap = cap;
    }
    return true;
}
}  // namespace

extern "C" bool DNX_shape_spheres(DNvolume* vol, const DNXsphereEdit* edits, uint32_t n)
{
    if(n == 0) return true;
    if(!vol || !edits)
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "DNX_shape_spheres: null volume or edit list");
        return false;
    }
    DNgpuVolume* g = vol->gpu;
    const uint32_t tiles = g->tiles;
    if(!g->dEditWords)
    {
        if(!dalloc(&g->dEditWords, (size_t)tiles * 512u, "device copy of the chunk words")) return false;
        g->editValid.assign(tiles, 0);
        g->editSlot.assign(tiles, -1);
    }

    // (tile, sphere) pairs in the order DN_shape_sphere visits them
    struct Pair { uint32_t tile, sphere; };
    std::vector<Pair> pairs;
    pairs.reserve((size_t)n * 64u);
    std::vector<DNeditSphere> sph(n);
    std::vector<float> off[3];          // per axis: each tile's nearest-voxel offset
    for(uint32_t e = 0; e < n; e++)
    {
        const DNXsphereEdit& in = edits[e];
        const SphereBox b = sphere_box(vol, in.voxel.material, in.flipNormals != 0, in.center, in.radius);
        DNeditSphere& d = sph[e];
        d.c[0] = in.center.x; d.c[1] = in.center.y; d.c[2] = in.center.z;
        d.r2 = b.r2; d.r12 = b.r12;
        d.material = in.voxel.material;
        d.flip = b.shell ? 1u : 0u;
        const DNcompressedVoxel cv = DN_compress_voxel(in.voxel);
        d.carveNormal = cv.normal;
        d.albedo = cv.albedo;
        const int lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
        const float cc[3] = {in.center.x, in.center.y, in.center.z};
        for(int a = 0; a < 3; a++)
        {
            off[a].clear();
            for(int m = lo[a]; m <= hi[a]; m++) off[a].push_back(nearest_offset(m * 8, cc[a]));
        }
        for(int mZ = b.lo.z; mZ <= b.hi.z; mZ++)
        for(int mY = b.lo.y; mY <= b.hi.y; mY++)
        for(int mX = b.lo.x; mX <= b.hi.x; mX++)
        {
            const float nx = off[0][mX - lo[0]], ny = off[1][mY - lo[1]], nz = off[2][mZ - lo[2]];
            if(!sphere_touches(b, nx * nx + ny * ny + nz * nz)) continue;
            pairs.push_back({flat(vol, DNivec3{mX, mY, mZ}), e});
        }
    }
    const uint32_t P = (uint32_t)pairs.size();
    if(P == 0) return true;

    // tiles in first-touch order, each with its pairs in edit order
    std::vector<int32_t>& slotOf = g->editSlot;
    std::vector<uint32_t> tileIds, pairCount;
    for(const Pair& q : pairs)
    {
        if(slotOf[q.tile] < 0)
        {
            slotOf[q.tile] = (int32_t)tileIds.size();
            tileIds.push_back(q.tile);
            pairCount.push_back(0);
        }
        pairCount[slotOf[q.tile]]++;
    }
    struct SlotReset {                  // every exit leaves the slot table clear
        std::vector<int32_t>& slot;
        const std::vector<uint32_t>& ids;
        ~SlotReset() { for(uint32_t t : ids) slot[t] = -1; }
    } slotReset{slotOf, tileIds};
    const uint32_t U = (uint32_t)tileIds.size();
    std::vector<uint32_t> stale;
    for(uint32_t u = 0; u < U; u++)
        if(vol->map[tileIds[u]].flag != 0 && !g->editValid[tileIds[u]]) stale.push_back(tileIds[u]);
    const uint32_t S = (uint32_t)stale.size();

    // staging layout: inputs, then outputs
    const size_t oSph = 0, oIds = align16(oSph + sizeof(DNeditSphere) * n), oInfo = align16(oIds + 4u * U),
                 oStart = align16(oInfo + 4u * U), oPSph = align16(oStart + 4u * (U + 1)),
                 oPSeq = align16(oPSph + 4u * P), oStale = align16(oPSeq + 4u * P),
                 oStaleW = align16(oStale + 4u * S), inBytes = align16(oStaleW + 4096u * S);
    const size_t oEv = inBytes, oOut = align16(oEv + P), oGW = align16(oOut + 4u * U),
                 allBytes = align16(oGW + 4096u * U);
    if(!ensure_edit_stage(g, allBytes, allBytes)) return false;
    if(!g->editEvents && !HIP_OK(hipEventCreateWithFlags(&g->editEvents, hipEventDisableTiming), "edit event"))
        return false;
    uint8_t* h = g->hEdit;
    uint8_t* d = g->dEdit;
    memcpy(h + oSph, sph.data(), sizeof(DNeditSphere) * n);
    memcpy(h + oIds, tileIds.data(), 4u * U);
    uint32_t* info = (uint32_t*)(h + oInfo);
    uint32_t* start = (uint32_t*)(h + oStart);
    start[0] = 0;
    for(uint32_t u = 0; u < U; u++)
    {
        const DNchunkHandle& hd = vol->map[tileIds[u]];
        info[u] = hd.flag ? (0x80000000u | vol->chunks[hd.chunkIndex].numVoxels) : 0u;
        start[u + 1] = start[u] + pairCount[u];
    }
    {
        uint32_t* pSph = (uint32_t*)(h + oPSph);
        uint32_t* pSeq = (uint32_t*)(h + oPSeq);
        std::vector<uint32_t> cur(start, start + U);
        for(uint32_t q = 0; q < P; q++)
        {
            const uint32_t k = cur[slotOf[pairs[q].tile]]++;
            pSph[k] = pairs[q].sphere;
            pSeq[k] = q;
        }
    }
    memcpy(h + oStale, stale.data(), 4u * S);
    for(uint32_t i = 0; i < S; i++)
        memcpy(h + oStaleW + 4096u * i, vol->chunks[vol->map[stale[i]].chunkIndex].voxels, 4096u);

    DNeditParams ep;
    ep.spheres = (const DNeditSphere*)(d + oSph);
    ep.tileIds = (const uint32_t*)(d + oIds);
    ep.tileInfo = (const uint32_t*)(d + oInfo);
    ep.pairStart = (const uint32_t*)(d + oStart);
    ep.pairSphere = (const uint32_t*)(d + oPSph);
    ep.pairSeq = (const uint32_t*)(d + oPSeq);
    ep.events = d + oEv;
    ep.tileOut = (uint32_t*)(d + oOut);
    ep.words = g->dEditWords;
    ep.mapX = vol->mapSize.x; ep.mapY = vol->mapSize.y; ep.mapZ = vol->mapSize.z;
    bool ok = HIP_OK(hipMemcpyAsync(d, h, inBytes, hipMemcpyHostToDevice, g->stream), "edit upload");
    ok = ok && dn_launch_edit_scatter(g->dEditWords, (const uint32_t*)(d + oStale), (const uint2*)(d + oStaleW), S,
                                      g->stream) == 0;
    ok = ok && dn_launch_sphere_edits(ep, U, g->stream) == 0;
    // the events and tile states come back first; the final words of every
    // touched tile follow while the host replays the events (copying them in
    // quarters, each copied into the CPU chunks once back, measured slower:
    // 0.88 against 0.82 ms per 256-edit call)
    ok = ok && HIP_OK(hipMemcpyAsync(h + oEv, d + oEv, oGW - oEv, hipMemcpyDeviceToHost, g->stream), "edit readback");
    ok = ok && HIP_OK(hipEventRecord(g->editEvents, g->stream), "edit event");
    ok = ok && dn_launch_edit_gather(g->dEditWords, ep.tileIds, (uint2*)(d + oGW), U, g->stream) == 0;
    ok = ok && HIP_OK(hipMemcpyAsync(h + oGW, d + oGW, 4096u * U, hipMemcpyDeviceToHost, g->stream), "edit readback");
    ok = ok && HIP_OK(hipEventSynchronize(g->editEvents), "sphere edits");
    if(!ok)
    {
        hipStreamSynchronize(g->stream);
        return false;
    }

    // Replay the chunk events in the reference's order. A removal clears the
    // chunk's 512 voxels (DN_remove_chunk does so at once); most slots freed
    // here are taken again by a later add of the same call, and every added
    // tile is written, so the copy-back below overwrites its whole chunk. The
    // voxels are therefore cleared at the end instead: those of slots still
    // free then, and on an early return those of every slot freed here.
    struct PendingClear {
        DNvolume* vol;
        std::vector<uint32_t> slots;
        bool copied;
        ~PendingClear()
        {
            for(uint32_t s : slots)
                if(!copied || !DN_in_map_bounds(vol, vol->chunks[s].pos)) clear_chunk_voxels(vol, s);
        }
    } pending{vol, {}, false};
    const uint8_t* ev = h + oEv;
    for(uint32_t q = 0; q < P; q++)
    {
        if(!ev[q]) continue;
        const uint32_t t = pairs[q].tile;
        if(!(ev[q] & (DN_EDIT_ADDED | DN_EDIT_REMOVED)))
        {
            mark_dirty(g, t);
            continue;
        }
        const DNivec3 mp = {(int)(t % vol->mapSize.x), (int)((t / vol->mapSize.x) % vol->mapSize.y),
                            (int)(t / (vol->mapSize.x * vol->mapSize.y))};
        if(ev[q] & DN_EDIT_ADDED)
        {
            if(DN_add_chunk(vol, mp) < 0)
            {
                dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_ERROR, "DNX_shape_spheres: no chunk memory for an edit");
                hipStreamSynchronize(g->stream);
                return false;
            }
        }
        else
        {
            // DN_remove_chunk with the voxel clearing left to `pending`
            const uint32_t mi = flat(vol, mp), slot = vol->map[mi].chunkIndex;
            vol->map[mi].flag = 0;
            vol->nextChunk = slot;
            clear_chunk_header(vol, slot);
            mark_dirty(g, mi);
            pending.slots.push_back(slot);
        }
    }
    // final words of the tiles that hold a chunk and were written
    const uint32_t* out = (const uint32_t*)(h + oOut);
    std::vector<uint32_t> gSlot;
    for(uint32_t u = 0; u < U; u++)
    {
        const bool present = (out[u] >> 31) != 0u, modified = ((out[u] >> 30) & 1u) != 0u;
        if(present != (vol->map[tileIds[u]].flag != 0))
        {
            dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "DNX_shape_spheres: chunk state mismatch at tile %u",
                   tileIds[u]);
            hipStreamSynchronize(g->stream);
            return false;
        }
        if(present && modified) gSlot.push_back(u);
    }
    if(!HIP_OK(hipStreamSynchronize(g->stream), "sphere edits")) return false;
    // copy the written chunks' words into the CPU chunks, split over a few
    // threads when there are many (each chunk is 4 KiB)
    const uint32_t G = (uint32_t)gSlot.size();
    auto copy_range = [&](uint32_t a, uint32_t b) {
        for(uint32_t i = a; i < b; i++)
        {
            const uint32_t u = gSlot[i];
            DNchunk& ch = vol->chunks[vol->map[tileIds[u]].chunkIndex];
            memcpy(ch.voxels, h + oGW + 4096u * u, 4096u);
            ch.numVoxels = out[u] & 0x3FFFFFFFu;
            ch.updated = true;
        }
    };
    const uint32_t workers = G >= 2048u ? 8u : (G >= 512u ? 4u : 1u);
    if(workers == 1u)
        copy_range(0, G);
    else
    {
        // a thread that cannot be started leaves its range to this one
        std::vector<std::thread> pool;
        const uint32_t per = (G + workers - 1u) / workers;
        uint32_t done = per;
        for(uint32_t w = 1; w < workers; w++)
        {
            try
            {
                pool.emplace_back(copy_range, std::min(G, w * per), std::min(G, (w + 1u) * per));
            }
            catch(const std::exception&)
            {
                break;
            }
            done = std::min(G, (w + 1u) * per);
        }
        copy_range(0, std::min(G, per));
        copy_range(done, G);
        for(std::thread& t : pool) t.join();
    }
    pending.copied = true;
    for(uint32_t u = 0; u < U; u++) g->editValid[tileIds[u]] = 1;
    return true;
}

// ===========================================================================
// DNX_* extensions

extern "C" void DNX_set_message_callback(void (*cb)(DNmessageType, DNmessageSeverity, const char*))
{
    g_DN_message_callback = cb;
}

extern "C" uint32_t DNX_error_count(void) { return g_errors.load(); }

extern "C" void DNX_set_voxels_dense(DNvolume* vol, DNivec3 origin, DNuvec3 size, const uint32_t* normalWords,
                                     const uint32_t* albedoWords)
{
    size_t k = 0;
    for(uint32_t z = 0; z < size.z; z++)
    for(uint32_t y = 0; y < size.y; y++)
    for(uint32_t x = 0; x < size.x; x++, k++)
    {
        DNivec3 p = {origin.x + (int32_t)x, origin.y + (int32_t)y, origin.z + (int32_t)z};
        if(p.x < 0 || p.y < 0 || p.z < 0) continue;
        DNivec3 mp = {p.x / 8, p.y / 8, p.z / 8}, cp = {p.x % 8, p.y % 8, p.z % 8};
        if(!DN_in_map_bounds(vol, mp)) continue;
        DNcompressedVoxel v = {normalWords[k], albedoWords ? albedoWords[k] : 0u};
        uint32_t mi = flat(vol, mp);
        if(vol->map[mi].flag == 0 && material_of(v.normal) == DN_MATERIAL_EMPTY) continue;
        DN_set_compressed_voxel(vol, mp, cp, v);
    }
}

extern "C" void DNX_get_voxels_dense(DNvolume* vol, DNivec3 origin, DNuvec3 size, uint32_t* normalWords,
                                     uint32_t* albedoWords)
{
    size_t k = 0;
    for(uint32_t z = 0; z < size.z; z++)
    for(uint32_t y = 0; y < size.y; y++)
    for(uint32_t x = 0; x < size.x; x++, k++)
    {
        DNivec3 p = {origin.x + (int32_t)x, origin.y + (int32_t)y, origin.z + (int32_t)z};
        normalWords[k] = UINT32_MAX;
        albedoWords[k] = 0u;
        if(p.x < 0 || p.y < 0 || p.z < 0) continue;
        DNivec3 mp = {p.x / 8, p.y / 8, p.z / 8}, cp = {p.x % 8, p.y % 8, p.z % 8};
        if(!DN_in_map_bounds(vol, mp) || vol->map[flat(vol, mp)].flag == 0) continue;
        DNcompressedVoxel v = DN_get_compressed_voxel(vol, mp, cp);
        normalWords[k] = v.normal;
        albedoWords[k] = v.albedo;
    }
}

extern "C" bool DNX_set_sky_cubemap(DNvolume* vol, const uint8_t* faces, uint32_t size)
{
    DNgpuVolume* g = vol->gpu;
    hipStreamSynchronize(g->stream);
    dfree(g->dCube);
    g->cubeSize = 0;
    if(!faces || size == 0) return true;
    if(size > 16384u)
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "cubemap faces of %u texels are too large", size);
        return false;
    }
    const size_t texels = (size_t)6 * size * size;
    std::vector<uchar4> rgba(texels);
    for(size_t i = 0; i < texels; i++) rgba[i] = make_uchar4(faces[3 * i], faces[3 * i + 1], faces[3 * i + 2], 255);
    if(!dalloc(&g->dCube, texels, "sky cubemap") ||
       !HIP_OK(hipMemcpy(g->dCube, rgba.data(), sizeof(uchar4) * texels, hipMemcpyHostToDevice), "sky cubemap upload"))
    {
        dfree(g->dCube);
        return false;
    }
    g->cubeSize = size;
    return true;
}

extern "C" void DNX_request_all_chunks(DNvolume* vol) { vol->gpu->requestAll = true; }

namespace {
// A voxel counts as see-through when it is empty, outside the map, or of a
// material with opacity < 1 (the culling rule of voxel.c:1391-1394, across chunks).
bool see_through(DNvolume* vol, int gx, int gy, int gz)
{
    if(gx < 0 || gy < 0 || gz < 0) return true;
    int cx = gx >> 3, cy = gy >> 3, cz = gz >> 3;
    if((uint32_t)cx >= vol->mapSize.x || (uint32_t)cy >= vol->mapSize.y || (uint32_t)cz >= vol->mapSize.z) return true;
    const DNchunkHandle& h = vol->map[(uint32_t)cx + vol->mapSize.x * ((uint32_t)cy + vol->mapSize.y * (uint32_t)cz)];
    if(h.flag == 0) return true;
    uint32_t m = material_of(vol->chunks[h.chunkIndex].voxels[gx & 7][gy & 7][gz & 7].normal);
    return m == DN_MATERIAL_EMPTY || vol->materials[m].opacity < 1.0f;
}

void compute_exposed(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    g->exposed.assign(g->tiles, 0);
    unsigned nth = std::min<unsigned>(8, std::max<unsigned>(1, std::thread::hardware_concurrency()));
    auto work = [&](unsigned w) {
        for(uint32_t t = w; t < g->tiles; t += nth)
        {
            if(vol->map[t].flag == 0) continue;
            const DNchunk& ch = vol->chunks[vol->map[t].chunkIndex];
            int bx = (int)(t % vol->mapSize.x) * 8, by = (int)((t / vol->mapSize.x) % vol->mapSize.y) * 8,
                bz = (int)(t / (vol->mapSize.x * vol->mapSize.y)) * 8;
            bool any = false;
            for(int x = 0; x < 8 && !any; x++)
            for(int y = 0; y < 8 && !any; y++)
            for(int z = 0; z < 8 && !any; z++)
            {
                if(material_of(ch.voxels[x][y][z].normal) == DN_MATERIAL_EMPTY) continue;
                int gx = bx + x, gy = by + y, gz = bz + z;
                any = see_through(vol, gx + 1, gy, gz) || see_through(vol, gx - 1, gy, gz) ||
                      see_through(vol, gx, gy + 1, gz) || see_through(vol, gx, gy - 1, gz) ||
                      see_through(vol, gx, gy, gz + 1) || see_through(vol, gx, gy, gz - 1);
            }
            g->exposed[t] = any ? 1 : 0;
        }
    };
    std::vector<std::thread> th;
    for(unsigned w = 0; w < nth; w++) th.emplace_back(work, w);
    for(auto& x : th) x.join();
    g->exposedValid = true;
}
}  // namespace

extern "C" uint32_t DNX_mark_visible(DNvolume* vol, uint32_t maxChunks)
{
    DNgpuVolume* g = vol->gpu;
    if(!g->exposedValid) { compute_exposed(vol); g->markValid = false; }
    if(!g->markValid || g->markMax != maxChunks)
    {
        std::vector<uint32_t> list;
        for(uint32_t t = 0; t < g->tiles && list.size() < maxChunks; t++)
            if(g->state[t] == 2 && g->exposed[t]) list.push_back(t);
        if(!g->dMarkList && !dalloc(&g->dMarkList, g->tiles, "visible list")) return 0;
        hipStreamSynchronize(g->stream);
        if(!list.empty() &&
           !HIP_OK(hipMemcpy(g->dMarkList, list.data(), sizeof(uint32_t) * list.size(), hipMemcpyHostToDevice),
                   "visible list upload"))
            return 0;
        g->markCount = (uint32_t)list.size();
        g->markMax = maxChunks;
        g->markValid = true;
    }
    if(g->markCount == 0) return 0;
    int e = dn_launch_mark_visible(g->dFlags, g->dMarkList, g->markCount, g->stream);
    if(e) dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "mark visible launch failed (%d)", e);
    return g->markCount;
}

extern "C" void* DNX_get_stream(DNvolume* vol) { return (void*)vol->gpu->stream; }

extern "C" bool DNX_synchronize(DNvolume* vol)
{
    return HIP_OK(hipStreamSynchronize(vol->gpu->stream), "stream synchronize");
}

extern "C" size_t DNX_fetch_lighting_requests(DNvolume* vol)
{
    DNgpuVolume* g = vol->gpu;
    if(!HIP_OK(hipMemcpyAsync(g->hSmall, g->dNumRequests, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream),
               "request readback") ||
       !HIP_OK(hipStreamSynchronize(g->stream), "request readback"))
        return 0;
    size_t n = g->hSmall[0];
    if(n > vol->lightingRequestCap && !DN_set_max_lighting_requests(vol, n)) return 0;
    if(n && !HIP_OK(hipMemcpy(vol->lightingRequests, g->dRequests, sizeof(uint32_t) * n, hipMemcpyDeviceToHost),
                    "request readback"))
        return 0;
    vol->numLightingRequests = n;
    return n;
}

extern "C" size_t DNX_export_gpu_state(DNvolume* vol, uint32_t* map, uint32_t* chunks, uint32_t* voxels,
                                       size_t voxelWords)
{
    DNgpuVolume* g = vol->gpu;
    if(!HIP_OK(hipStreamSynchronize(g->stream), "export")) return 0;
    const uint32_t T = g->tiles;
    if(map || chunks)
    {
        std::vector<uint32_t> fl(T), lu(T);
        std::vector<DNchunkRec> recs(T);
        if(!HIP_OK(hipMemcpy(fl.data(), g->dFlags, sizeof(uint32_t) * T, hipMemcpyDeviceToHost), "export") ||
           !HIP_OK(hipMemcpy(lu.data(), g->dLastUsed, sizeof(uint32_t) * T, hipMemcpyDeviceToHost), "export") ||
           !HIP_OK(hipMemcpy(recs.data(), g->dChunks, sizeof(DNchunkRec) * T, hipMemcpyDeviceToHost), "export"))
            return 0;
        for(uint32_t t = 0; t < T; t++)
        {
            bool loaded = (fl[t] & 3u) == 2u;
            if(map)
            {
                map[3 * t] = fl[t];
                map[3 * t + 1] = lu[t];
                map[3 * t + 2] = loaded ? recs[t].voxelIndex : 0u;
            }
            if(chunks)
            {
                uint32_t* c = chunks + (size_t)24 * t;
                memset(c, 0, 24 * sizeof(uint32_t));
                if(loaded)
                {
                    c[0] = (uint32_t)recs[t].pos[0]; c[1] = (uint32_t)recs[t].pos[1]; c[2] = (uint32_t)recs[t].pos[2];
                    c[3] = recs[t].samples;
                    c[4] = recs[t].partial[0]; c[5] = recs[t].partial[1]; c[6] = recs[t].partial[2];
                    memcpy(c + 7, recs[t].mask, 16 * sizeof(uint32_t));
                }
            }
        }
    }
    if(voxels && voxelWords)
    {
        size_t words = std::min(voxelWords, (size_t)g->pool.cap * 4);
        if(!HIP_OK(hipMemcpy(voxels, g->dVoxels, sizeof(uint32_t) * words, hipMemcpyDeviceToHost), "export"))
            return 0;
    }
    return g->pool.cap;
}

extern "C" uint32_t DNX_get_ray_points(DNvolume* vol, int which, float* out, uint32_t cap)
{
    std::vector<float> pts = fibonacci_points(which == 0 ? vol->numSpecularRays : vol->numDirectRays);
    uint32_t n = (uint32_t)(pts.size() / 3);
    if(n > cap) n = cap;
    memcpy(out, pts.data(), sizeof(float) * 3 * n);
    return n;
}

extern "C" void DNX_get_counters(DNvolume* vol, int kind, DNXcounters* out, bool reset)
{
    DNgpuVolume* g = vol->gpu;
    if(kind < 0 || kind > 2) { memset(out, 0, sizeof(*out)); return; }
    uint64_t c[CNT_SLOTS];
    if(!HIP_OK(hipStreamSynchronize(g->stream), "counters") ||
       !HIP_OK(hipMemcpy(c, g->dCounters + kind * CNT_SLOTS, sizeof(c), hipMemcpyDeviceToHost), "counters"))
    {
        memset(out, 0, sizeof(*out));
        return;
    }
    out->rays = c[CNT_RAYS];
    out->mapSteps = c[CNT_MAPSTEPS];
    out->voxelSteps = c[CNT_VOXSTEPS];
    out->hits = c[CNT_HITS];
    out->litVoxels = c[CNT_LIT];
    out->pixels = c[CNT_PIXELS];
    if(reset)
    {
        hipMemset(g->dCounters + kind * CNT_SLOTS, 0, sizeof(c));
        hipDeviceSynchronize();
    }
}

extern "C" void DNX_enable_timing(DNvolume* vol, bool enable) { vol->gpu->timing = enable; }

extern "C" void DNX_get_timing(DNvolume* vol, DNXtiming* out, bool reset)
{
    DNgpuVolume* g = vol->gpu;
    timers_collect(g);
    out->drawMs = g->ms[TK_DRAW]; out->lightingMs = g->ms[TK_LIGHT]; out->visibleMs = g->ms[TK_VISIBLE];
    out->uploadMs = g->ms[TK_UPLOAD]; out->epilogueMs = g->ms[TK_EPILOGUE]; out->raysMs = g->ms[TK_RAYS];
    out->drawLaunches = g->launches[TK_DRAW]; out->lightingLaunches = g->launches[TK_LIGHT];
    out->visibleLaunches = g->launches[TK_VISIBLE]; out->uploadLaunches = g->launches[TK_UPLOAD];
    out->epilogueLaunches = g->launches[TK_EPILOGUE]; out->raysLaunches = g->launches[TK_RAYS];
    if(reset)
        for(int i = 0; i < 6; i++) { g->ms[i] = 0; g->launches[i] = 0; }
}

extern "C" void DNX_cast_rays(DNvolume* vol, uint32_t n, const float* origins, const float* dirs, bool ignoreFirst,
                              DNXrayHit* out)
{
    DNgpuVolume* g = vol->gpu;
    DNrayParams p;
    p.scene = scene_view(vol);
    p.n = n;
    p.origins = origins;
    p.dirs = dirs;
    p.ignoreFirst = ignoreFirst ? 1u : 0u;
    p.out = out;
    p.counters = g->dCounters + 2 * CNT_SLOTS;
    if(!upload_materials(vol)) return;
    Timer tm = timer_begin(g, TK_RAYS);
    int e = dn_launch_rays(p, g->stream);
    timer_end(g, tm);
    if(e) dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR, "ray launch failed (%d)", e);
}

extern "C" void DNX_set_draw_hit_output(DNvolume* vol, DNXpixelHit* out, size_t capacityPixels)
{
    vol->gpu->drawHitOut = out;
    vol->gpu->drawHitCap = out ? capacityPixels : 0;
}

extern "C" void DNX_set_lighting_debug_output(DNvolume* vol, float* values, uint32_t* voxelIndices,
                                              size_t requestCapacity)
{
    vol->gpu->dbgValues = values;
    vol->gpu->dbgIndex = voxelIndices;
    vol->gpu->dbgCap = requestCapacity;
}

extern "C" void DNX_get_upload_stats(DNvolume* vol, uint64_t* bytes, uint64_t* chunks, uint64_t* voxels, bool reset)
{
    DNgpuVolume* g = vol->gpu;
    if(bytes) *bytes = g->upBytes;
    if(chunks) *chunks = g->upChunks;
    if(voxels) *voxels = g->upVoxels;
    if(reset) g->upBytes = g->upChunks = g->upVoxels = 0;
}

extern "C" void DNX_set_lighting_sample_output(DNvolume* vol, float* samples, size_t requestCapacity)
{
    vol->gpu->freshOut = samples;
    vol->gpu->freshCap = samples ? requestCapacity : 0;
}

// Sizes of the public structs, so bindings can check their mirrors.
extern "C" size_t DNX_struct_size(int which)
{
    switch(which)
    {
    case 0: return sizeof(DNvolume);
    case 1: return sizeof(DNchunk);
    case 2: return sizeof(DNmaterial);
    case 3: return sizeof(DNchunkHandle);
    case 4: return sizeof(DNvoxelNode);
    case 5: return sizeof(DNXcounters);
    case 6: return sizeof(DNXtiming);
    case 7: return sizeof(DNXrayHit);
    case 8: return sizeof(DNXpixelHit);
    case 9: return sizeof(DNvoxel);
    case 10: return offsetof(DNvolume, numSpecularRays);
    case 11: return offsetof(DNvolume, camPos);
    case 12: return sizeof(DNXbatchConfig);
    case 13: return sizeof(DNXrcConfig);
    case 14: return sizeof(DNXsphereEdit);
    default: return 0;
    }
}

// ===========================================================================
// BATCHED LIGHTING (no counterpart in the reference)

struct DNXbatch {
    std::vector<DNvolume*> vols;
    hipStream_t stream = nullptr;
    DNXbatchConfig cfg = {8, 256, 0, 1};
    DNlightParams* hParams[2] = {nullptr, nullptr};   // pinned, one per frame slot
    DNlightParams* dParams = nullptr;
    uint32_t* dQueueCounters = nullptr;
    uint32_t* dQueueScenes = nullptr;                 // [8][n]
    uint32_t* dQueueLen = nullptr;                    // [8]
    uint32_t* dQueuePrefix = nullptr;                 // [8][n + 1]
    uint32_t queues = 0;                              // queue tables uploaded for this many queues
    hipEvent_t ready = nullptr;                       // re-recorded on each volume's stream
    hipEvent_t slotDone[2] = {nullptr, nullptr};
    bool slotUsed[2] = {false, false};
    int slot = 0;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    DNXbatchConfig execCfg[2];
    uint32_t execTiles[2] = {0, 0}, execReq[2] = {0, 0};
    bool execDirect[2] = {false, false};
    bool anyDirect = false;                           // some volume traces its direct rays ahead this frame
    bool timing = true;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, spare;
    double ms = 0.0;
    uint32_t launches = 0;
};

namespace {

bool batch_config_ok(const DNXbatchConfig& c)
{
    return c.requestsPerUnit >= 2u && c.requestsPerUnit <= 64u && !(c.requestsPerUnit & 1u) &&
           (c.blockThreads == 64u || c.blockThreads == 256u) && c.blocksPerCU <= 64u && c.xcdQueues <= 1u;
}

// Deals the volumes to the queues (volume i -> queue i % Q, in order) and
// uploads the tables; synchronous, outside any capture.
bool batch_upload_queues(DNXbatch* b)
{
    const uint32_t n = (uint32_t)b->vols.size();
    const uint32_t Q = b->cfg.xcdQueues ? (uint32_t)DN_BATCH_MAX_QUEUES : 1u;
    if(b->queues == Q) return true;
    hipStreamSynchronize(b->stream);   // an earlier launch may still read the tables
    std::vector<uint32_t> ids((size_t)DN_BATCH_MAX_QUEUES * n, 0u), len(DN_BATCH_MAX_QUEUES, 0u);
    for(uint32_t i = 0; i < n; i++)
    {
        uint32_t q = i % Q;
        ids[(size_t)q * n + len[q]++] = i;
    }
    if(!HIP_OK(hipMemcpy(b->dQueueScenes, ids.data(), sizeof(uint32_t) * ids.size(), hipMemcpyHostToDevice),
               "batch queue table") ||
       !HIP_OK(hipMemcpy(b->dQueueLen, len.data(), sizeof(uint32_t) * len.size(), hipMemcpyHostToDevice),
               "batch queue table"))
        return false;
    b->queues = Q;
    return true;
}

DNbatchParams batch_params(DNXbatch* b, bool dry)
{
    DNbatchParams bp;
    bp.scenes = b->dParams;
    bp.numScenes = (uint32_t)b->vols.size();
    bp.unit = b->cfg.requestsPerUnit;
    bp.queues = b->queues;
    bp.dry = dry ? 1u : 0u;
    bp.queueCounters = b->dQueueCounters;
    bp.directCounters = b->dQueueCounters + DN_BATCH_MAX_QUEUES;
    bp.queueScenes = b->dQueueScenes;
    bp.queueLen = b->dQueueLen;
    bp.queuePrefix = b->dQueuePrefix;
    bp.dryLastUsed = nullptr;
    return bp;
}

// Fills the parameter slot for every volume and orders the batch stream after
// each volume's stream.
bool batch_prepare(DNXbatch* b, DNlightParams* dst, int nds, int mds, float time, uint32_t& maxTiles,
                   uint32_t& maxReq)
{
    maxTiles = 0;
    maxReq = 0;
    b->anyDirect = false;
    for(size_t i = 0; i < b->vols.size(); i++)
    {
        DNvolume* vol = b->vols[i];
        uint32_t mr = 0;
        if(!prepare_lighting(vol, nds, mds, time, dst[i], mr)) return false;
        DNgpuVolume* g = vol->gpu;
        if(dst[i].numDirectRays > 0 && dst[i].numDiffuseSamples > 0 && mr > 0 && ensure_direct(g, mr))
        {
            dst[i].directOut = g->dDirect;
            dst[i].directIn = g->dDirect;
            dst[i].directSlots = (uint32_t)std::min<size_t>(g->directCap, 0xFFFFFFFFu);
            b->anyDirect = true;
        }
        maxTiles = std::max(maxTiles, vol->gpu->tiles);
        maxReq = std::max(maxReq, mr);
        // the batch stream waits for the work queued on the volume's stream (its
        // sync's uploads and visible pass); a stream with none left needs no wait,
        // which saves two runtime calls per volume and frame (8.2 -> 3.1 ms of host
        // time per call for 384 volumes)
        if(hipStreamQuery(vol->gpu->stream) != hipSuccess)
        {
            hipEventRecord(b->ready, vol->gpu->stream);
            hipStreamWaitEvent(b->stream, b->ready, 0);
        }
    }
    return true;
}

void batch_free_execs(DNXbatch* b)
{
    for(int s = 0; s < 2; s++)
        if(b->exec[s])
        {
            hipGraphExecDestroy(b->exec[s]);
            b->exec[s] = nullptr;
        }
}

}  // namespace

extern "C" DNXbatch* DNX_batch_create(DNvolume* const* volumes, uint32_t count)
{
    if(!volumes || count == 0 || count > 65535u)
    {
        dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_ERROR, "batch: 1..65535 volumes required");
        return nullptr;
    }
    for(uint32_t i = 0; i < count; i++)
        if(!volumes[i] || !volumes[i]->gpu)
        {
            dn_msg(DN_MESSAGE_CPU_MEMORY, DN_MESSAGE_ERROR, "batch: volume %u is not a live volume", i);
            return nullptr;
        }
    DNXbatch* b = new DNXbatch();
    b->vols.assign(volumes, volumes + count);
    bool ok = HIP_OK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "batch stream") &&
              HIP_OK(hipEventCreateWithFlags(&b->ready, hipEventDisableTiming), "batch event");
    for(int s = 0; ok && s < 2; s++)
        ok = HIP_OK(hipHostMalloc((void**)&b->hParams[s], sizeof(DNlightParams) * count, hipHostMallocDefault),
                    "batch parameters") &&
             HIP_OK(hipEventCreateWithFlags(&b->slotDone[s], hipEventDisableTiming), "batch event");
    ok = ok && dalloc(&b->dParams, count, "batch parameters") &&
         dalloc(&b->dQueueCounters, 2 * DN_BATCH_MAX_QUEUES, "batch queues") &&
         dalloc(&b->dQueueScenes, (size_t)DN_BATCH_MAX_QUEUES * count, "batch queues") &&
         dalloc(&b->dQueueLen, DN_BATCH_MAX_QUEUES, "batch queues") &&
         dalloc(&b->dQueuePrefix, (size_t)DN_BATCH_MAX_QUEUES * (count + 1u), "batch queues");
    if(!ok)
    {
        DNX_batch_destroy(b);
        return nullptr;
    }
    return b;
}

extern "C" void DNX_batch_destroy(DNXbatch* b)
{
    if(!b) return;
    if(b->stream) hipStreamSynchronize(b->stream);
    batch_free_execs(b);
    for(auto& t : b->pending) { hipEventDestroy(t.first); hipEventDestroy(t.second); }
    for(auto& t : b->spare) { hipEventDestroy(t.first); hipEventDestroy(t.second); }
    for(int s = 0; s < 2; s++)
    {
        if(b->hParams[s]) hipHostFree(b->hParams[s]);
        if(b->slotDone[s]) hipEventDestroy(b->slotDone[s]);
    }
    dfree(b->dParams); dfree(b->dQueueCounters); dfree(b->dQueueScenes); dfree(b->dQueueLen); dfree(b->dQueuePrefix);
    if(b->ready) hipEventDestroy(b->ready);
    if(b->stream) hipStreamDestroy(b->stream);
    delete b;
}

extern "C" bool DNX_batch_set_config(DNXbatch* b, const DNXbatchConfig* c)
{
    if(!b || !c || !batch_config_ok(*c))
    {
        dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR,
               "batch config: requestsPerUnit even in [2, 64], blockThreads 64 or 256, blocksPerCU <= 64, "
               "xcdQueues 0 or 1");
        return false;
    }
    b->cfg = *c;
    return true;
}

extern "C" void DNX_batch_get_config(DNXbatch* b, DNXbatchConfig* c)
{
    if(b && c) *c = b->cfg;
}

extern "C" bool DNX_batch_update_lighting(DNXbatch* b, int numDiffuseSamples, int maxDiffuseSamples, float time)
{
    if(!b) return false;
    const int s = b->slot;
    if(b->slotUsed[s]) hipEventSynchronize(b->slotDone[s]);   // the slot's parameters were consumed
    if(!batch_upload_queues(b)) return false;
    uint32_t maxTiles = 0, maxReq = 0;
    if(!batch_prepare(b, b->hParams[s], numDiffuseSamples, maxDiffuseSamples, time, maxTiles, maxReq)) return false;
    const DNXbatchConfig& c = b->cfg;
    const DNXbatchConfig& e = b->execCfg[s];
    bool stale = !b->exec[s] || e.requestsPerUnit != c.requestsPerUnit || e.blockThreads != c.blockThreads ||
                 e.blocksPerCU != c.blocksPerCU || e.xcdQueues != c.xcdQueues || b->execTiles[s] != maxTiles ||
                 b->execReq[s] != maxReq || b->execDirect[s] != b->anyDirect;
    if(stale)
    {
        if(b->exec[s]) { hipGraphExecDestroy(b->exec[s]); b->exec[s] = nullptr; }
        // capture on a private stream so the waits on the volume streams stay outside the graph
        hipStream_t cs = nullptr;
        if(!HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "batch capture stream")) return false;
        hipGraph_t graph = nullptr;
        bool ok = HIP_OK(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed), "batch capture");
        if(ok)
        {
            const bool copied = HIP_OK(hipMemcpyAsync(b->dParams, b->hParams[s], sizeof(DNlightParams) * b->vols.size(),
                                                      hipMemcpyHostToDevice, cs), "batch parameters");
            int le = dn_launch_batch(batch_params(b, false), c.blockThreads, c.blocksPerCU, maxTiles, maxReq, true,
                                     b->anyDirect, cs);
            // the capture is ended whatever failed inside it
            ok = HIP_OK(hipStreamEndCapture(cs, &graph), "batch capture") && le == 0 && copied;
            if(ok) ok = HIP_OK(hipGraphInstantiate(&b->exec[s], graph, nullptr, nullptr, 0), "batch graph");
            if(graph) hipGraphDestroy(graph);
        }
        hipStreamDestroy(cs);
        if(!ok) return false;
        b->execCfg[s] = c;
        b->execTiles[s] = maxTiles;
        b->execReq[s] = maxReq;
        b->execDirect[s] = b->anyDirect;
    }
    std::pair<hipEvent_t, hipEvent_t> t(nullptr, nullptr);
    if(b->timing)
    {
        if(!b->spare.empty()) { t = b->spare.back(); b->spare.pop_back(); }
        else { hipEventCreate(&t.first); hipEventCreate(&t.second); }
        hipEventRecord(t.first, b->stream);
    }
    bool ok = HIP_OK(hipGraphLaunch(b->exec[s], b->stream), "batch launch");
    if(t.first)
    {
        hipEventRecord(t.second, b->stream);
        b->pending.push_back(t);
    }
    hipEventRecord(b->slotDone[s], b->stream);
    b->slotUsed[s] = true;
    for(DNvolume* vol : b->vols) hipStreamWaitEvent(vol->gpu->stream, b->slotDone[s], 0);
    b->slot = s ^ 1;
    return ok;
}

extern "C" double DNX_batch_timing(DNXbatch* b, uint32_t* launches, bool reset)
{
    if(!b) return 0.0;
    for(auto& t : b->pending)
    {
        hipEventSynchronize(t.second);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, t.first, t.second);
        b->ms += ms;
        b->launches++;
        b->spare.push_back(t);
    }
    b->pending.clear();
    double ms = b->ms;
    if(launches) *launches = b->launches;
    if(reset) { b->ms = 0.0; b->launches = 0; }
    return ms;
}

extern "C" bool DNX_batch_autotune(DNXbatch* b, int numDiffuseSamples, int maxDiffuseSamples, float time, int repeats,
                                   DNXbatchConfig* best, float* bestMs)
{
    if(!b) return false;
    if(repeats < 1) repeats = 1;
    for(int s = 0; s < 2; s++)
        if(b->slotUsed[s]) hipEventSynchronize(b->slotDone[s]);
    uint32_t maxTiles = 0, maxReq = 0;
    DNlightParams* hp = b->hParams[b->slot];
    if(!batch_prepare(b, hp, numDiffuseSamples, maxDiffuseSamples, time, maxTiles, maxReq)) return false;
    if(!HIP_OK(hipMemcpyAsync(b->dParams, hp, sizeof(DNlightParams) * b->vols.size(), hipMemcpyHostToDevice,
                              b->stream), "batch parameters"))
        return false;
    uint32_t* scratchAges = nullptr;            // the dry passes' lastUsed stores
    if(!dalloc(&scratchAges, std::max<uint32_t>(maxTiles, 1u), "autotune scratch")) return false;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const DNXbatchConfig keep = b->cfg;
    DNXbatchConfig win = keep;
    float winMs = -1.0f;
    const uint32_t units[] = {2, 4, 8, 16, 32};
    const uint32_t threads[] = {64, 256};
    bool ok = true;
    for(uint32_t xq = 0; xq < 2 && ok; xq++)
        for(uint32_t th : threads)
            for(uint32_t u : units)
            {
                b->cfg = DNXbatchConfig{u, th, 0, xq};
                if(!batch_upload_queues(b)) { ok = false; break; }
                float ms = 0.0f, mn = -1.0f;
                for(int r = 0; r < repeats; r++)
                {
                    hipEventRecord(e0, b->stream);
                    DNbatchParams bp = batch_params(b, true);
                    bp.dryLastUsed = scratchAges;
                    dn_launch_batch(bp, th, 0, maxTiles, maxReq, false, b->anyDirect, b->stream);
                    hipEventRecord(e1, b->stream);
                    hipEventSynchronize(e1);
                    hipEventElapsedTime(&ms, e0, e1);
                    if(mn < 0.0f || ms < mn) mn = ms;
                }
                if(winMs < 0.0f || mn < winMs) { winMs = mn; win = b->cfg; }
            }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamSynchronize(b->stream);
    dfree(scratchAges);
    b->cfg = ok ? win : keep;
    ok = ok && batch_upload_queues(b);
    hipEventRecord(b->ready, b->stream);
    for(DNvolume* vol : b->vols) hipStreamWaitEvent(vol->gpu->stream, b->ready, 0);
    if(best) *best = b->cfg;
    if(bestMs) *bestMs = winMs;
    return ok;
}

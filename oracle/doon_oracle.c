/*
 * doon_oracle.c -- CPU restatement of DoonEngine's GPU voxel pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke
 * check and the benchmark's CPU-baseline leg load it. The product path never
 * links or calls it.
 *
 * It restates, operation by operation, the reference's GLSL compute shaders and
 * the host-side pieces of voxel.c that feed them:
 *   assets/shaders/voxelShared.comp   layouts, codec, rank/select, two-level DDA
 *   assets/shaders/voxelDraw.comp     primary-ray draw cast
 *   assets/shaders/voxelLighting.comp per-voxel lighting
 *   src/DoonEngine/voxel.c:1391-1461  _DN_check_face_visible / _DN_chunk_to_gpu
 *   src/DoonEngine/voxel.c:1463-1489  _DN_request_chunk_lighting
 *
 * The same source is compiled twice:
 *   default      FP32, strict IEEE (no contraction, no fast-math): the bit-exact
 *                pin for the HIP kernels.
 *   -DOR_DOUBLE  FP64 variant of the same algorithm: the accuracy oracle.
 *
 * GLSL built-ins are restated with fixed semantics (GLSL leaves several of them
 * implementation-defined):
 *   min(x,y) = y<x ? y : x       max(x,y) = x<y ? y : x     clamp = min(max(x,lo),hi)
 *   sign(x)  = x>0 ? 1 : x<0 ? -1 : 0                        fract(x) = x - floor(x)
 *   mix(x,y,a) = x*(1-a) + y*a   dot = (x*x' + y*y') + z*z'
 *   normalize(v) = v * (1/sqrt(dot(v,v)))                    reflect = I - (2*dot(N,I))*N
 *   round = round-half-to-even   sin = C library sinf / sin  pow = powf / pow
 *
 * Deterministic semantics replace the reference's data races (voxelLighting.comp
 * :224/:284-285 sample count, :104-105/:281 visible bit, :113/:122 neighbour
 * lighting reads): every read during a lighting pass sees the state at dispatch
 * start, and the pass's writes go to separate output buffers; visible bits set
 * by specular rays are merged after the lit chunks' bits are cleared.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef OR_DOUBLE
typedef double real;
#define RL(x) (x)
#define r_floor floor
#define r_sqrt sqrt
#define r_sin sin
#define r_pow pow
#define r_trunc trunc
#define r_rint nearbyint
#define r_fabs fabs
#else
typedef float real;
#define RL(x) (x##f)
#define r_floor floorf
#define r_sqrt sqrtf
#define r_sin sinf
#define r_pow powf
#define r_trunc truncf
#define r_rint nearbyintf
#define r_fabs fabsf
#endif

#define EPS RL(0.0001)               /* voxelShared.comp:7 */
#define INV255 RL(0.00392156862)     /* voxelShared.comp:249-253 (FP32: one ulp below 1/255) */
#define INV65535 RL(0.0000152590219) /* voxelShared.comp:252 */

/* ------------------------------------------------------------------------- */
/* vector helpers with the GLSL semantics listed above                        */

typedef struct { real x, y, z; } v3;
typedef struct { int32_t x, y, z; } iv3;

static inline v3 V(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 vs(real s) { return V(s, s, s); }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vmuls(v3 a, real s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vadds(v3 a, real s) { return V(a.x + s, a.y + s, a.z + s); }
static inline v3 vdivs(v3 a, real s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline real vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline real gmin(real x, real y) { return y < x ? y : x; }
static inline real gmax(real x, real y) { return x < y ? y : x; }
static inline real gsign(real x) { return x > 0 ? RL(1.0) : (x < 0 ? RL(-1.0) : RL(0.0)); }
static inline v3 vfloor(v3 a) { return V(r_floor(a.x), r_floor(a.y), r_floor(a.z)); }
static inline v3 vabs(v3 a) { return V(r_fabs(a.x), r_fabs(a.y), r_fabs(a.z)); }
static inline v3 vsign(v3 a) { return V(gsign(a.x), gsign(a.y), gsign(a.z)); }
static inline v3 vinv(v3 a) { return V(RL(1.0) / a.x, RL(1.0) / a.y, RL(1.0) / a.z); }
static inline v3 vmin(v3 a, v3 b) { return V(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return V(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
static inline v3 vclamp(v3 a, real lo, real hi)
{
    return V(gmin(gmax(a.x, lo), hi), gmin(gmax(a.y, lo), hi), gmin(gmax(a.z, lo), hi));
}
static inline v3 vnorm(v3 a)
{
    real inv = RL(1.0) / r_sqrt(vdot(a, a));
    return vmuls(a, inv);
}
static inline v3 vreflect(v3 I, v3 N) { return vsub(I, vmuls(N, RL(2.0) * vdot(N, I))); }
static inline v3 vrefract(v3 I, v3 N, real eta)
{
    real d = vdot(N, I);
    real k = RL(1.0) - eta * eta * (RL(1.0) - d * d);
    if(k < 0) return vs(0);
    return vsub(vmuls(I, eta), vmuls(N, eta * d + r_sqrt(k)));
}
static inline v3 vmix(v3 x, v3 y, real a) { return vadd(vmuls(x, RL(1.0) - a), vmuls(y, a)); }
static inline v3 vfromi(iv3 a) { return V((real)a.x, (real)a.y, (real)a.z); }
static inline v3 vf(const float* p) { return V((real)p[0], (real)p[1], (real)p[2]); }
static inline real fractr(real x) { return x - r_floor(x); }
static inline real minc(v3 a) { return gmin(gmin(a.x, a.y), a.z); }
static inline v3 vpow(v3 a, real e) { return V(r_pow(a.x, e), r_pow(a.y, e), r_pow(a.z, e)); }

/* ------------------------------------------------------------------------- */
/* reference buffer layouts (voxelShared.comp:19-95, voxel.c:16-42)            */
/* map tile : [0] flags [1] lastUsed [2] voxelIndex                            */
/* chunk    : [0..2] pos [3] numIndirectSamples [4..6] partialCounts           */
/*            [7..22] bitMask [23] padding                                     */
/* voxel    : [0] material<<24|normal [1] albedo<<8|spec.x                     */
/*            [2] spec.y<<24|spec.z<<16|diff.x [3] diff.y<<16|diff.z           */

enum { TILE_WORDS = 3, CHUNK_WORDS = 24, VOX_WORDS = 4 };

typedef struct ORmaterial {       /* voxel.h:82-94, std140 32 B */
    float padding[2];             /* padding[0] > 0: emission strength (0 means 1.0) */
    uint32_t emissive;
    float opacity;
    float refractIndex;
    float specular;
    uint32_t reflectType;
    uint32_t shininess;
} ORmaterial;

typedef struct ORscene {
    const uint32_t* map;          /* tiles * 3 words: snapshot at dispatch start */
    const uint32_t* chunks;       /* tiles * 24 words */
    const uint32_t* voxels;       /* pool * 4 words */
    const ORmaterial* mats;       /* 256 */
    uint32_t mapSize[3];
    float sunStrength[3];
    float skyBot[3], skyTop[3];
    uint32_t* flagsOut;           /* tiles words: receives request writes (flags = 3), may be NULL */
    uint32_t* lastUsedOut;        /* tiles words: receives lastUsed = 0, may be NULL */
    const uint8_t* cube;          /* sky cubemap: 6 faces (+X, -X, +Y, -Y, +Z, -Z) of cubeSize^2 RGB8, or NULL */
    uint32_t cubeSize;
} ORscene;

typedef struct ORcounters {
    uint64_t calls;   /* step_map invocations */
    uint64_t sMap;    /* chunk-grid loop trips (voxelShared.comp:435) */
    uint64_t sVox;    /* voxel loop trips (voxelShared.comp:339) */
    uint64_t hits;    /* step_map calls that returned true */
} ORcounters;

typedef struct ORvoxel { v3 normal; uint32_t material; v3 albedo; v3 spec; v3 diffuse; } ORvoxel;

/* per-invocation shader globals (voxelShared.comp:321-325) plus instrumentation */
typedef struct ORthread {
    int enableRefraction;
    uint32_t lastVoxID;
    real lastVoxRefract;
    uint32_t hitIndex;            /* voxel pool index of the last hit */
    uint32_t sMap, sVox;          /* trips of the last step_map call */
    ORcounters cnt;
    int stopOutside;              /* lighting rays: stop beyond the box of present tiles (below) */
    int32_t presLo[3], presHi[3]; /* that box, inclusive; lo > hi when no tile is present */
    const int32_t* colTop;        /* column tables of present_column_tops, or NULL */
} ORthread;

/* The GPU's lighting rays end a walk once it is beyond the box of present
 * tiles (map flags & 3 != 0) on the side it moves toward: past presHi on an
 * axis it steps up, below presLo on any other (dn_device.hpp, step_map). No
 * tile there is entered, requested or stamped, so the walk would end as the
 * same miss with the same colour terms; only the trip counters differ. The
 * lighting pass applies the rule while this is set (the default), so the
 * GPU's counters can be compared; tests clear it to show that the rule
 * changes nothing else. */
int or_stop_outside = 1;

static void present_box(const ORscene* s, int32_t lo[3], int32_t hi[3])
{
    for(int a = 0; a < 3; a++) { lo[a] = INT32_MAX; hi[a] = INT32_MIN; }
    for(uint32_t z = 0; z < s->mapSize[2]; z++)
        for(uint32_t y = 0; y < s->mapSize[1]; y++)
            for(uint32_t x = 0; x < s->mapSize[0]; x++)
            {
                size_t t = (size_t)x + s->mapSize[0] * ((size_t)y + (size_t)s->mapSize[1] * z);
                if(!(s->map[t * TILE_WORDS] & 3u)) continue;
                int32_t c[3] = {(int32_t)x, (int32_t)y, (int32_t)z};
                for(int a = 0; a < 3; a++)
                {
                    if(c[a] < lo[a]) lo[a] = c[a];
                    if(c[a] > hi[a]) hi[a] = c[a];
                }
            }
}

/* The second rule: per column (x, z) the highest present tile, then for each
 * quadrant q (bit 0: x steps up, else x stays or steps down; bit 1: the same
 * for z) its maximum over the columns at or beyond (x, z) in those directions:
 * out[q][z][x], -1 where there is none. A lighting ray whose y does not step
 * down ends once its y after a step is above the value of the cell it left
 * (the GPU's DNsceneView::colTop and step_map). out holds 4 * X * Z values;
 * each quadrant is filled by the recurrence M(x, z) = max(top(x, z),
 * M(x + dx, z), M(x, z + dz)), visiting the cells from the far corner. */
void or_present_column_tops(const ORscene* s, int32_t* out)
{
    const int32_t X = (int32_t)s->mapSize[0], Y = (int32_t)s->mapSize[1], Z = (int32_t)s->mapSize[2];
    const size_t plane = (size_t)X * (size_t)Z;
    int32_t* top = (int32_t*)malloc((plane ? plane : 1) * sizeof(int32_t));
    for(size_t i = 0; i < plane; i++) top[i] = -1;
    for(int32_t z = 0; z < Z; z++)
        for(int32_t y = 0; y < Y; y++)
            for(int32_t x = 0; x < X; x++)
                if(s->map[((size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z)) * TILE_WORDS] & 3u)
                    top[(size_t)z * X + x] = y;
    for(int q = 0; q < 4; q++)
    {
        int32_t* m = out + (size_t)q * plane;
        const int32_t dx = (q & 1) ? 1 : -1, dz = (q & 2) ? 1 : -1;
        for(int32_t iz = 0; iz < Z; iz++)
        {
            const int32_t z = dz > 0 ? Z - 1 - iz : iz;
            for(int32_t ix = 0; ix < X; ix++)
            {
                const int32_t x = dx > 0 ? X - 1 - ix : ix;
                int32_t v = top[(size_t)z * X + x];
                if(x + dx >= 0 && x + dx < X && m[(size_t)z * X + (x + dx)] > v) v = m[(size_t)z * X + (x + dx)];
                if(z + dz >= 0 && z + dz < Z && m[(size_t)(z + dz) * X + x] > v) v = m[(size_t)(z + dz) * X + x];
                m[(size_t)z * X + x] = v;
            }
        }
    }
    free(top);
}

static void thread_init(ORthread* t, int refraction)
{
    memset(t, 0, sizeof(*t));
    t->enableRefraction = refraction;
    t->lastVoxID = 255u;
    t->lastVoxRefract = RL(1.0);
    t->hitIndex = 0xFFFFFFFFu;
}

/* decompress_voxel, voxelShared.comp:236-256 */
static ORvoxel decompress(const uint32_t* w)
{
    ORvoxel r;
    uint32_t n0 = w[0] >> 24, n1 = (w[0] >> 16) & 255u, n2 = (w[0] >> 8) & 255u, n3 = w[0] & 255u;
    uint32_t a0 = w[1] >> 24, a1 = (w[1] >> 16) & 255u, a2 = (w[1] >> 8) & 255u, a3 = w[1] & 255u;
    uint32_t s0 = w[2] >> 24, s1 = (w[2] >> 16) & 255u, s2 = (w[2] >> 8) & 255u, s3 = w[2] & 255u;
    uint32_t d0 = w[3] >> 24, d1 = (w[3] >> 16) & 255u, d2 = (w[3] >> 8) & 255u, d3 = w[3] & 255u;
    uint32_t dx = (s2 << 8) | s3, dy = (d0 << 8) | d1, dz = (d2 << 8) | d3;
    r.normal = vmuls(vsub(vmuls(V((real)n1, (real)n2, (real)n3), INV255), vs(RL(0.5))), RL(2.0));
    r.material = n0;
    r.albedo = vmuls(V((real)a0, (real)a1, (real)a2), INV255);
    r.diffuse = vmuls(V((real)dx, (real)dy, (real)dz), INV65535);
    r.spec = vmuls(V((real)a3, (real)s0, (real)s1), INV255);
    return r;
}

static inline real emission(const ORmaterial* m) { return m->padding[0] > 0.0f ? (real)m->padding[0] : RL(1.0); }

/* ------------------------------------------------------------------------- */
/* occupancy rank / select (voxelShared.comp:150-231)                          */

static inline uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

/* get_voxel_index without the voxelIndex base: stored voxels before (x,y,z) */
uint32_t or_rank(const uint32_t* chunk, uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t li = x + 8u * (y + 8u * z);
    uint32_t w = li >> 5;
    const uint32_t* mask = chunk + 7;
    uint32_t n = (w > 3) ? chunk[4 + (w >> 2) - 1] : 0u;
    for(uint32_t i = w & ~3u; i <= w; i++)
    {
        uint32_t bits = mask[i];
        if(i == w) bits &= (1u << (li & 31u)) - 1u;
        n += popc(bits);
    }
    return n;
}

/* get_voxel_position: position index x + 8y + 64z of the (voxNum+1)-th stored
 * voxel, or -1 when the chunk stores fewer voxels (voxelShared.comp:172-224). */
int32_t or_select(const uint32_t* chunk, uint32_t voxNum)
{
    const uint32_t* mask = chunk + 7;
    uint32_t count = 0, pos = 0, start = 0;
    for(int i = 0; i < 3; i++)
    {
        if(chunk[4 + i] < voxNum) { count = chunk[4 + i]; pos += 128; start += 4; }
        else break;
    }
    voxNum++;
    for(uint32_t i = start; i < 16; i++)
    {
        if(count == voxNum) break;
        uint32_t bc = popc(mask[i]);
        if(count + bc >= voxNum)
        {
            uint32_t bit = 0;
            while(count < voxNum) { count += (mask[i] >> bit) & 1u; bit++; pos++; }
        }
        else { count += bc; pos += 32; }
    }
    pos--;
    if(count < voxNum) return -1;
    return (int32_t)pos;
}

static inline int voxel_exists(const uint32_t* chunk, iv3 p)
{
    uint32_t li = (uint32_t)(p.x + 8 * (p.y + 8 * p.z));
    return (chunk[7 + (li >> 5)] >> (li & 31u)) & 1u;
}

/* ------------------------------------------------------------------------- */
/* two-level DDA (voxelShared.comp:300-475)                                    */

typedef struct { iv3 pos; v3 delta; iv3 step; v3 side; } dda_t;

static void init_dda(v3 dir, v3 inv, v3 p, dda_t* d)
{
    v3 f = vfloor(p);
    d->pos.x = (int32_t)f.x; d->pos.y = (int32_t)f.y; d->pos.z = (int32_t)f.z;
    d->delta = vabs(inv);
    v3 s = vsign(dir);
    d->step.x = (int32_t)s.x; d->step.y = (int32_t)s.y; d->step.z = (int32_t)s.z;
    d->side = vmul(vadds(vadd(vmul(s, vsub(vfromi(d->pos), p)), vmuls(s, RL(0.5))), RL(0.5)), d->delta);
}

static void iterate_dda(dda_t* d, v3* normal)
{
    v3 sd = d->side;
    int mx = sd.x <= gmin(sd.y, sd.z);
    int my = sd.y <= gmin(sd.z, sd.x);
    int mz = sd.z <= gmin(sd.x, sd.y);
    real fx = mx ? RL(1.0) : RL(0.0), fy = my ? RL(1.0) : RL(0.0), fz = mz ? RL(1.0) : RL(0.0);
    d->side.x = sd.x + fx * d->delta.x;
    d->side.y = sd.y + fy * d->delta.y;
    d->side.z = sd.z + fz * d->delta.z;
    d->pos.x += mx * d->step.x; d->pos.y += my * d->step.y; d->pos.z += mz * d->step.z;
    normal->x = fx * (real)(-d->step.x);
    normal->y = fy * (real)(-d->step.y);
    normal->z = fz * (real)(-d->step.z);
}

static inline int in_chunk(iv3 p) { return p.x < 8 && p.y < 8 && p.z < 8 && p.x >= 0 && p.y >= 0 && p.z >= 0; }
static inline int in_map(const ORscene* s, iv3 p)
{
    return p.x >= 0 && p.y >= 0 && p.z >= 0 && (uint32_t)p.x < s->mapSize[0] && (uint32_t)p.y < s->mapSize[1] &&
           (uint32_t)p.z < s->mapSize[2];
}
static inline uint32_t map_index_u(const ORscene* s, uint32_t x, uint32_t y, uint32_t z)
{
    return x + s->mapSize[0] * (y + s->mapSize[1] * z);
}

/* step_chunk, voxelShared.comp:328-418 (maxDepth is negative on every path here,
 * so the depth test at :363 always passes) */
static int step_chunk(const ORscene* s, ORthread* t, uint32_t mapIndex, v3* dir, v3* inv, v3* p, int ignoreFirst,
                      v3* hitNormal, ORvoxel* vox, v3* colorAdd, real* colorMult, int* refracted)
{
    *refracted = 0;
    const uint32_t* chunk = s->chunks + (size_t)mapIndex * CHUNK_WORDS;
    dda_t d;
    init_dda(*dir, *inv, *p, &d);
    v3 lsd = vs(0);
    while(in_chunk(d.pos))
    {
        t->sVox++;
        if(voxel_exists(chunk, d.pos) && !ignoreFirst)
        {
            uint32_t idx = s->map[(size_t)mapIndex * TILE_WORDS + 2] +
                           or_rank(chunk, (uint32_t)d.pos.x, (uint32_t)d.pos.y, (uint32_t)d.pos.z);
            const uint32_t* cw = s->voxels + (size_t)idx * VOX_WORDS;
            *vox = decompress(cw);
            const ORmaterial* m = &s->mats[vox->material];
            uint32_t thisID = (cw[1] & 0xFFFFFF00u) | (cw[0] >> 24);
            if(m->opacity == 1.0f)
            {
                *p = vadd(*p, vmuls(*dir, minc(lsd) + EPS));
                t->hitIndex = idx;
                return 1;
            }
            else if(t->lastVoxID != thisID)
            {
                real op = (real)m->opacity;
                *colorAdd = vadd(*colorAdd, vmul(vmuls(vox->albedo, *colorMult * op), vf(s->sunStrength)));
                *colorMult = *colorMult * (RL(1.0) - op);
                real rayDist = minc(lsd);
                if(t->enableRefraction && rayDist > 0)
                {
                    *refracted = 1;
                    *p = vadd(*p, vmuls(*dir, rayDist + EPS));
                    v3 n = vdot(vox->normal, *dir) < 0 ? vnorm(vox->normal) : *hitNormal;
                    *dir = vrefract(*dir, n, t->lastVoxRefract / (real)m->refractIndex);
                    *inv = vinv(*dir);
                    init_dda(*dir, *inv, *p, &d);
                    lsd = vs(0);
                }
                t->lastVoxID = thisID;
                t->lastVoxRefract = (real)m->refractIndex;
            }
        }
        else if(t->lastVoxID != 255u)
        {
            if(t->enableRefraction)
            {
                *refracted = 1;
                *p = vadd(*p, vmuls(*dir, minc(lsd) + EPS));
                v3 old = *dir;
                v3 nn = vneg(vox->normal);
                v3 n = vdot(nn, *dir) < 0 ? vnorm(nn) : *hitNormal;
                *dir = vrefract(*dir, n, t->lastVoxRefract);
                if(dir->x == 0 && dir->y == 0 && dir->z == 0) *dir = old;
                *inv = vinv(*dir);
                init_dda(*dir, *inv, *p, &d);
                lsd = vs(0);
            }
            t->lastVoxID = 255u;
            t->lastVoxRefract = RL(1.0);
        }
        lsd = d.side;
        iterate_dda(&d, hitNormal);
        ignoreFirst = 0;
    }
    return 0;
}

/* step_map, voxelShared.comp:421-475. *inv is the caller's 1/dir; refraction
 * may update dir (inout) and the local inverse. */
static int step_map(const ORscene* s, ORthread* t, v3* dir, v3 inv, v3* p, int ignoreFirst, v3* hitNormal,
                    ORvoxel* vox, v3* colorAdd, real* colorMult)
{
    int requested = 0;
    *colorAdd = vs(0);
    *colorMult = RL(1.0);
    t->sMap = 0; t->sVox = 0; t->hitIndex = 0xFFFFFFFFu;
    dda_t d;
    init_dda(*dir, inv, *p, &d);
    v3 lsd = vs(0);
    int hit = 0;
    /* the column table of the ray's quadrant, for a ray whose y does not step down */
    const int32_t* ct = (t->stopOutside && t->colTop && d.step.y >= 0)
                            ? t->colTop + (size_t)((d.step.x > 0 ? 1 : 0) + (d.step.z > 0 ? 2 : 0)) *
                                              s->mapSize[0] * s->mapSize[2]
                            : NULL;
    while(in_map(s, d.pos))
    {
        t->sMap++;
        /* highest present tile over the columns reachable from this cell */
        const int32_t reach = ct ? ct[(size_t)d.pos.z * s->mapSize[0] + (size_t)d.pos.x] : INT32_MAX;
        uint32_t idx = map_index_u(s, (uint32_t)d.pos.x, (uint32_t)d.pos.y, (uint32_t)d.pos.z);
        uint32_t flags = s->map[(size_t)idx * TILE_WORDS];
        if((flags & 3u) == 2u)
        {
            if(s->lastUsedOut) s->lastUsedOut[idx] = 0;
            v3 upd = vadd(*p, vmuls(*dir, minc(lsd) - EPS));
            v3 crp = vmuls(vsub(upd, vfromi(d.pos)), RL(8.0));
            crp = vclamp(crp, EPS, RL(8.0) - EPS);
            int refr;
            if(step_chunk(s, t, idx, dir, &inv, &crp, ignoreFirst, hitNormal, vox, colorAdd, colorMult, &refr))
            {
                *p = vadd(vfromi(d.pos), vmuls(crp, RL(0.125)));
                hit = 1;
                break;
            }
            if(refr)
            {
                *p = vadd(vfromi(d.pos), vmuls(crp, RL(0.125)));
                init_dda(*dir, inv, *p, &d);
                lsd = vs(0);
            }
        }
        else if(!requested && (flags & 3u) != 0)
        {
            if(s->flagsOut) s->flagsOut[idx] = 3u;
            requested = 1;
        }
        lsd = d.side;
        iterate_dda(&d, hitNormal);
        ignoreFirst = 0;
        if(t->stopOutside && ((d.step.x > 0 ? d.pos.x > t->presHi[0] : d.pos.x < t->presLo[0]) ||
                              (d.step.y > 0 ? d.pos.y > t->presHi[1] : d.pos.y < t->presLo[1]) ||
                              (d.step.z > 0 ? d.pos.z > t->presHi[2] : d.pos.z < t->presLo[2]) ||
                              d.pos.y > reach))
            break;
    }
    t->cnt.calls++;
    t->cnt.sMap += t->sMap;
    t->cnt.sVox += t->sVox;
    t->cnt.hits += (uint64_t)hit;
    return hit;
}

/* texture(samplerCube, dir) for the reference's sky cubemap (main.c:335-353:
 * GL_LINEAR, GL_CLAMP_TO_EDGE, RGB8 faces uploaded as GL_TEXTURE_CUBE_MAP_POSITIVE_X
 * + i). Cube map face selection after the OpenGL specification's table
 * (major axis ma; +X: sc = -z, tc = -y; -X: sc = z, tc = -y; +Y: sc = x,
 * tc = z; -Y: sc = x, tc = -z; +Z: sc = x, tc = -y; -Z: sc = -x, tc = -y;
 * s = (sc/|ma| + 1)/2, t = (tc/|ma| + 1)/2), ties to x, then y (the
 * specification leaves ties to the implementation); GL_LINEAR over texel
 * centres, u = s*n - 1/2, weights (1-a)(1-b), a(1-b), (1-a)b, ab, indices
 * clamped to the face. UNORM8 texels read as c/255. */
static v3 cube_texel(const uint8_t* face, uint32_t n, int32_t i, int32_t j)
{
    i = i < 0 ? 0 : (i > (int32_t)n - 1 ? (int32_t)n - 1 : i);
    j = j < 0 ? 0 : (j > (int32_t)n - 1 ? (int32_t)n - 1 : j);
    const uint8_t* t = face + 3u * ((uint32_t)j * n + (uint32_t)i);
    return V((real)t[0] / RL(255.0), (real)t[1] / RL(255.0), (real)t[2] / RL(255.0));
}

static v3 cube_sample(const uint8_t* cube, uint32_t n, v3 d)
{
    real ax = r_fabs(d.x), ay = r_fabs(d.y), az = r_fabs(d.z), sc, tc, ma;
    uint32_t f;
    if(ax >= ay && ax >= az) { f = d.x >= 0 ? 0u : 1u; sc = d.x >= 0 ? -d.z : d.z; tc = -d.y; ma = ax; }
    else if(ay >= az) { f = d.y >= 0 ? 2u : 3u; sc = d.x; tc = d.y >= 0 ? d.z : -d.z; ma = ay; }
    else { f = d.z >= 0 ? 4u : 5u; sc = d.z >= 0 ? d.x : -d.x; tc = -d.y; ma = az; }
    real s = (sc / ma + RL(1.0)) * RL(0.5), t = (tc / ma + RL(1.0)) * RL(0.5);
    real u = s * (real)n - RL(0.5), v = t * (real)n - RL(0.5);
    real fu = r_floor(u), fv = r_floor(v), a = u - fu, b = v - fv;
    int32_t i0 = (int32_t)fu, j0 = (int32_t)fv;
    const uint8_t* face = cube + (size_t)f * n * n * 3u;
    v3 c00 = cube_texel(face, n, i0, j0), c10 = cube_texel(face, n, i0 + 1, j0);
    v3 c01 = cube_texel(face, n, i0, j0 + 1), c11 = cube_texel(face, n, i0 + 1, j0 + 1);
    return vadd(vadd(vadd(vmuls(c00, (RL(1.0) - a) * (RL(1.0) - b)), vmuls(c10, a * (RL(1.0) - b))),
                     vmuls(c01, (RL(1.0) - a) * b)), vmuls(c11, a * b));
}

/* sky_color, voxelShared.comp:289-295 */
static inline v3 sky_color(const ORscene* s, v3 dir)
{
    if(s->cube) return cube_sample(s->cube, s->cubeSize, dir);
    return vmix(vf(s->skyBot), vf(s->skyTop), dir.y * RL(0.5) + RL(1.0));
}

/* sky_color alone, for tests: out = rgb */
void or_sky_color(const ORscene* s, const float* dir, real* out)
{
    v3 c = sky_color(s, vf(dir));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

/* ------------------------------------------------------------------------- */
/* single-ray entry (tests): one step_map call in map units                    */

typedef struct ORrayHit {
    uint32_t hit, voxelIndex, sMap, sVox;
    float pos[3];
    float normal[3];
} ORrayHit;

void or_cast_rays(const ORscene* s, uint32_t n, const float* origins, const float* dirs, uint32_t ignoreFirst,
                  ORrayHit* out, ORcounters* cnt)
{
    ORthread t;
    thread_init(&t, 0);
    for(uint32_t i = 0; i < n; i++)
    {
        v3 dir = vf(dirs + 3 * i), p = vf(origins + 3 * i), nrm = vs(0), ca;
        real cm;
        ORvoxel vox;
        memset(&vox, 0, sizeof(vox));
        t.lastVoxID = 255u; t.lastVoxRefract = RL(1.0);
        int h = step_map(s, &t, &dir, vinv(dir), &p, (int)ignoreFirst, &nrm, &vox, &ca, &cm);
        out[i].hit = (uint32_t)h;
        out[i].voxelIndex = h ? t.hitIndex : 0xFFFFFFFFu;
        out[i].sMap = t.sMap; out[i].sVox = t.sVox;
        out[i].pos[0] = (float)p.x; out[i].pos[1] = (float)p.y; out[i].pos[2] = (float)p.z;
        out[i].normal[0] = (float)nrm.x; out[i].normal[1] = (float)nrm.y; out[i].normal[2] = (float)nrm.z;
    }
    if(cnt) *cnt = t.cnt;
}

/* ------------------------------------------------------------------------- */
/* draw cast (voxelDraw.comp:24-147). Raster compositing and the cubemap sky   */
/* are not on this path (composeRasterized false, useCubemap false).           */

typedef struct ORdrawParams {
    float invView[16];        /* column-major m[col][row] (QuickMath layout) */
    float invViewProj[16];    /* invCenteredView * invProjection, multiplied on the host */
    uint32_t width, height;
    uint32_t viewMode;
} ORdrawParams;

typedef struct ORpixel {
    uint32_t voxelIndex, sMap, sVox, pad;
    float face[3];
} ORpixel;

static v3 voxel_color(const ORscene* s, uint32_t viewMode, ORvoxel vox, v3 ca, real cm, v3 hitNormal)
{
    const ORmaterial* m = &s->mats[vox.material];
    vox.spec = vmuls(vox.spec, (real)m->specular);
    vox.diffuse = vmuls(vox.diffuse, RL(1.0) - (real)m->specular);
    switch(viewMode)
    {
    case 0: {
        v3 solid = m->emissive ? vmuls(vox.albedo, emission(m)) : vadd(vmul(vox.diffuse, vox.albedo), vox.spec);
        return vadd(vmuls(solid, cm), ca);
    }
    case 1: return vox.albedo;
    case 2: return m->emissive ? vox.albedo : vox.diffuse;
    case 3: return m->emissive ? vox.albedo : vox.spec;
    case 4: return vabs(vox.normal);
    default: return vabs(hitNormal);
    }
}

static void draw_rows(const ORscene* s, const ORdrawParams* dp, uint32_t y0, uint32_t y1, float* rgba, ORpixel* px,
                      uint32_t* flagsOut, ORcounters* cnt)
{
    ORthread t;
    thread_init(&t, 1);
    const float* M = dp->invViewProj;
    v3 ms = V((real)s->mapSize[0], (real)s->mapSize[1], (real)s->mapSize[2]);
    for(uint32_t y = y0; y < y1; y++)
    for(uint32_t x = 0; x < dp->width; x++)
    {
        t.lastVoxID = 255u; t.lastVoxRefract = RL(1.0);
        size_t o = (size_t)y * dp->width + x;
        real sx = (real)x / (real)dp->width * RL(2.0) - RL(1.0);
        real sy = (real)y / (real)dp->height * RL(2.0) - RL(1.0);
        v3 org = V((real)dp->invView[12], (real)dp->invView[13], (real)dp->invView[14]);
        v3 p = org;
        v3 q = V((real)M[0] * sx + (real)M[4] * sy + (real)M[8] * RL(0.0) + (real)M[12],
                 (real)M[1] * sx + (real)M[5] * sy + (real)M[9] * RL(0.0) + (real)M[13],
                 (real)M[2] * sx + (real)M[6] * sy + (real)M[10] * RL(0.0) + (real)M[14]);
        v3 dir = vadds(vnorm(q), EPS);
        v3 inv = vinv(dir);
        v3 color;
        real depth = RL(-1.0);
        ORpixel info = {0xFFFFFFFFu, 0, 0, 0, {0, 0, 0}};
        v3 tMin = vmul(vsub(vs(0), p), inv), tMax = vmul(vsub(ms, p), inv);
        v3 t1 = vmin(tMin, tMax), t2 = vmax(tMin, tMax);
        real tNear = gmax(gmax(t1.x, t1.y), t1.z), tFar = gmin(gmin(t2.x, t2.y), t2.z);
        if(tNear > tFar || tFar < 0)
            color = vpow(sky_color(s, dir), RL(2.2));
        else
        {
            if(tNear > 0) p = vadd(p, vmuls(dir, tNear + EPS));
            /* normal_AABB, voxelShared.comp:276-284 */
            v3 c = vmuls(vadd(vs(0), ms), RL(0.5));
            v3 pc = vsub(p, c);
            v3 dd = vmuls(vsub(ms, vs(0)), RL(0.5));
            real bias = RL(1.0) + EPS;
            v3 tr = V(r_trunc(pc.x / dd.x * bias), r_trunc(pc.y / dd.y * bias), r_trunc(pc.z / dd.z * bias));
            v3 nrm = vnorm(tr);
            ORvoxel vox;
            memset(&vox, 0, sizeof(vox));
            v3 ca; real cm;
            int hit = step_map(s, &t, &dir, inv, &p, 0, &nrm, &vox, &ca, &cm);
            info.sMap = t.sMap; info.sVox = t.sVox;
            if(hit)
            {
                v3 fp = vfloor(p);
                if(flagsOut && fp.x >= 0 && fp.y >= 0 && fp.z >= 0 && fp.x < ms.x && fp.y < ms.y && fp.z < ms.z)
                    flagsOut[map_index_u(s, (uint32_t)fp.x, (uint32_t)fp.y, (uint32_t)fp.z)] |= 4u;
                v3 dv = vsub(p, org);
                depth = r_sqrt(vdot(dv, dv));
                color = voxel_color(s, dp->viewMode, vox, ca, cm, nrm);
                info.voxelIndex = t.hitIndex;
            }
            else
                color = vadd(vmuls(vpow(sky_color(s, dir), RL(2.2)), cm), ca);
            info.face[0] = (float)nrm.x; info.face[1] = (float)nrm.y; info.face[2] = (float)nrm.z;
        }
        color = vpow(color, RL(0.4545));
        rgba[4 * o + 0] = (float)color.x; rgba[4 * o + 1] = (float)color.y;
        rgba[4 * o + 2] = (float)color.z; rgba[4 * o + 3] = (float)depth;
        if(px) px[o] = info;
    }
    if(cnt) *cnt = t.cnt;
}


typedef struct DRange {
    const ORscene* s;
    const ORdrawParams* dp;
    uint32_t y0, y1;
    float* rgba;
    ORpixel* px;
    uint32_t* visible;      /* this range's visible-bit writes (a copy of flagsOut) */
    uint32_t* requests;     /* this range's request writes (a copy of the scene's flagsOut) */
    ORcounters cnt;
} DRange;

static void* draw_rows_thread(void* arg)
{
    DRange* j = (DRange*)arg;
    ORscene sc = *j->s;
    sc.flagsOut = j->requests;
    sc.lastUsedOut = NULL;
    draw_rows(&sc, j->dp, j->y0, j->y1, j->rgba, j->px, j->visible, &j->cnt);
    return NULL;
}

/* The draw cast split into `numThreads` bands of rows run on host threads.
 * Pixels are independent; the visible bits (flagsOut |= 4) and ray requests
 * (the scene's flagsOut = 3) of the bands are merged afterwards, so the result
 * equals or_draw's. lastUsed writes are not collected. */
void or_draw_mt(const ORscene* s, const ORdrawParams* dp, float* rgba, ORpixel* px, uint32_t* flagsOut, ORcounters* cnt,
                uint32_t numThreads)
{
    uint32_t tiles = s->mapSize[0] * s->mapSize[1] * s->mapSize[2];
    size_t tw = tiles ? tiles : 1;
    if(numThreads < 1) numThreads = 1;
    if(numThreads > dp->height) numThreads = dp->height ? dp->height : 1;
    DRange* jobs = (DRange*)calloc(numThreads, sizeof(DRange));
    pthread_t* th = (pthread_t*)calloc(numThreads, sizeof(pthread_t));
    for(uint32_t k = 0; k < numThreads; k++)
    {
        DRange* j = &jobs[k];
        j->s = s; j->dp = dp; j->rgba = rgba; j->px = px;
        j->y0 = (uint32_t)((uint64_t)dp->height * k / numThreads);
        j->y1 = (uint32_t)((uint64_t)dp->height * (k + 1) / numThreads);
        if(flagsOut)
        {
            j->visible = (uint32_t*)malloc(sizeof(uint32_t) * tw);
            memcpy(j->visible, flagsOut, sizeof(uint32_t) * tiles);
        }
        if(s->flagsOut)
        {
            j->requests = (uint32_t*)malloc(sizeof(uint32_t) * tw);
            memcpy(j->requests, s->flagsOut, sizeof(uint32_t) * tiles);
        }
        pthread_create(&th[k], NULL, draw_rows_thread, j);
    }
    for(uint32_t k = 0; k < numThreads; k++) pthread_join(th[k], NULL);
    ORcounters total = {0, 0, 0, 0};
    for(uint32_t k = 0; k < numThreads; k++)
    {
        DRange* j = &jobs[k];
        total.calls += j->cnt.calls; total.sMap += j->cnt.sMap; total.sVox += j->cnt.sVox; total.hits += j->cnt.hits;
        for(uint32_t i = 0; flagsOut && i < tiles; i++) flagsOut[i] |= j->visible[i] & 4u;
        for(uint32_t i = 0; s->flagsOut && i < tiles; i++)
            if(j->requests[i] != s->flagsOut[i] && j->requests[i] == 3u) s->flagsOut[i] = 3u;
        free(j->visible); free(j->requests);
    }
    free(th);
    free(jobs);
    if(cnt) *cnt = total;
}

void or_draw(const ORscene* s, const ORdrawParams* dp, float* rgba, ORpixel* px, uint32_t* flagsOut, ORcounters* cnt)
{
    draw_rows(s, dp, 0, dp->height, rgba, px, flagsOut, cnt);
}

/* ------------------------------------------------------------------------- */
/* hash RNG (voxelLighting.comp:29-57)                                         */

/* Optional override of the hash for given seeds (tests only): lets a check feed
 * the values another implementation's sin produced for exactly these seeds. */
static const float* g_ovSeeds = NULL;
static const float* g_ovValues = NULL;
static uint32_t g_ovCount = 0;

void or_set_rand_override(const float* seeds, const float* values, uint32_t n)
{
    g_ovSeeds = seeds; g_ovValues = values; g_ovCount = seeds && values ? n : 0;
}

static inline real rnd(real seed)
{
    if(g_ovCount)
    {
        float fs = (float)seed;
        uint32_t lo = 0, hi = g_ovCount;   /* seeds sorted ascending */
        while(lo < hi)
        {
            uint32_t mid = (lo + hi) / 2;
            if(g_ovSeeds[mid] < fs) lo = mid + 1; else hi = mid;
        }
        if(lo < g_ovCount && g_ovSeeds[lo] == fs && (real)fs == seed) return (real)g_ovValues[lo];
    }
    return fractr(r_sin(seed) * RL(43758.5453)) * RL(2.0) - RL(1.0);
}
static inline v3 rnd3(real seed) { return V(rnd(seed), rnd(seed * RL(2.0)), rnd(seed * RL(3.0))); }
#define RUS_MAX_TRIES 256   /* the reference loop is unbounded; this cap is never reached in practice */
static v3 rand_unit_sphere(real seed)
{
    v3 p = vs(0);
    for(int k = 0; k < RUS_MAX_TRIES; k++)
    {
        p = rnd3(seed);
        seed = seed + RL(1.0);
        if(vdot(p, p) >= RL(1.0)) continue;
        return p;
    }
    return p;
}

float or_rand(float seed) { return (float)rnd((real)seed); }
void or_rand_unit_sphere(float seed, float* out)
{
    v3 p = rand_unit_sphere((real)seed);
    out[0] = (float)p.x; out[1] = (float)p.y; out[2] = (float)p.z;
}

/* ------------------------------------------------------------------------- */
/* lighting pass (voxelLighting.comp:65-288)                                   */

typedef struct ORlightParams {
    float time;
    uint32_t numDiffuseSamples, maxDiffuseSamples, diffuseBounceLimit, specBounceLimit;
    float sunDir[3];              /* normalised on the host (voxel.c:942) */
    float ambient[3];
    float shadowSoftness;
    float camPos[3];
    uint32_t numSpecRays;         /* 15 reproduces the reference's spherePoints */
    uint32_t numDirectRays;       /* 0: one jittered shadow ray per sample (reference) */
    const float* specPoints;      /* numSpecRays * 3 */
    const float* directPoints;    /* numDirectRays * 3 */
} ORlightParams;

typedef struct LCtx {
    const ORscene* s;
    const ORlightParams* lp;
    ORthread* t;
    int firstSample;
    v3 sunDir, sunStrength;
    uint8_t* visProp;
} LCtx;

static void shadow_ray(LCtx* c, v3 p, real seed, v3* color)
{
    const ORlightParams* lp = c->lp;
    ORvoxel hv; v3 ca, tn = vs(0); real cm;
    memset(&hv, 0, sizeof(hv));
    if(lp->numDirectRays == 0)
    {
        v3 dir = c->firstSample ? vadds(c->sunDir, EPS)
                                : vadds(vnorm(vadd(vmuls(c->sunDir, (real)lp->shadowSoftness), rand_unit_sphere(seed))), EPS);
        v3 pp = p;
        if(!step_map(c->s, c->t, &dir, vinv(dir), &pp, 1, &tn, &hv, &ca, &cm))
            *color = vadd(*color, vadd(vmuls(c->sunStrength, cm), ca));
        return;
    }
    v3 acc = vs(0);
    for(uint32_t k = 0; k < lp->numDirectRays; k++)
    {
        v3 dir = vadds(vnorm(vadd(vmuls(c->sunDir, (real)lp->shadowSoftness), vf(lp->directPoints + 3 * k))), EPS);
        v3 pp = p;
        if(!step_map(c->s, c->t, &dir, vinv(dir), &pp, 1, &tn, &hv, &ca, &cm))
            acc = vadd(acc, vadd(vmuls(c->sunStrength, cm), ca));
    }
    *color = vadd(*color, vdivs(acc, (real)lp->numDirectRays));
}

static void specular_ray(LCtx* c, uint32_t thisMapIndex, v3 p, v3 dir, v3 albedo, uint32_t reflectType, v3* color)
{
    const ORscene* s = c->s;
    ORvoxel hv; v3 tn = vs(0);
    memset(&hv, 0, sizeof(hv));
    v3 lastPos = p;
    v3 mult = albedo;
    for(uint32_t i = 0; i < c->lp->specBounceLimit; i++)
    {
        v3 ca; real cm;
        if(step_map(s, c->t, &dir, vinv(dir), &p, 1, &tn, &hv, &ca, &cm))
        {
            /* ivec3(rayPos) truncates toward zero */
            int32_t hx = (int32_t)p.x, hy = (int32_t)p.y, hz = (int32_t)p.z;
            iv3 hp = {hx, hy, hz};
            if(in_map(s, hp) && (s->map[(size_t)thisMapIndex * TILE_WORDS] & 4u) && c->visProp)
                c->visProp[map_index_u(s, (uint32_t)hx, (uint32_t)hy, (uint32_t)hz)] = 1;
            v3 dist = vabs(vsub(vfloor(vmuls(p, RL(8.0))), vfloor(vmuls(lastPos, RL(8.0)))));
            if(vdot(dist, dist) <= RL(1.0)) return;
            const ORmaterial* hm = &s->mats[hv.material];
            hv.diffuse = vmuls(hv.diffuse, RL(1.0) - (real)hm->specular);
            if(hm->emissive)
            {
                *color = vadd(*color, vmul(vmul(vadd(vmuls(vmuls(hv.albedo, emission(hm)), cm), ca), mult), albedo));
                return;
            }
            v3 hitColor = vmul(hv.diffuse, hv.albedo);
            *color = vadd(*color, vmul(vadd(vmuls(hitColor, cm), ca), mult));
            if(hm->specular == 0.0f) return;
            mult = vmul(mult, vmuls(vmuls(hv.albedo, cm), (real)hm->specular));
            reflectType = hm->reflectType;
            lastPos = p;
            dir = vreflect(dir, hv.normal);
        }
        else if(vdot(dir, c->sunDir) > RL(0.99))
        {
            *color = vadd(*color, vadd(vmuls(c->sunStrength, cm), ca));
            return;
        }
        else
        {
            v3 src = reflectType == 1u ? sky_color(s, dir) : c->sunStrength;
            *color = vadd(*color, vmul(vadd(vmuls(src, cm), ca), mult));
            return;
        }
    }
}

static void diffuse_ray(LCtx* c, v3 normal, v3 p, ORvoxel initial, real seed, v3* color)
{
    const ORscene* s = c->s;
    const ORlightParams* lp = c->lp;
    ORvoxel hv = initial;
    v3 hitNormal = normal;
    const ORmaterial* hm = NULL;
    v3 newColor = vs(RL(1.0));
    v3 lastPos = p;
    v3 lastDir = vs(0);
    for(uint32_t i = 0; i < lp->diffuseBounceLimit; i++)
    {
        v3 dir;
        if(i > 0 && (rnd(seed + (real)lp->diffuseBounceLimit + (real)i) + RL(1.0)) * RL(0.5) < (real)hm->specular)
            dir = vnorm(vadd(vmuls(vreflect(lastDir, hitNormal), (real)hm->shininess),
                             rand_unit_sphere((real)lp->time + (real)i)));
        else if(c->firstSample)
            dir = vadds(vnorm(hitNormal), EPS);
        else
            dir = vadds(vnorm(vadd(hitNormal, rand_unit_sphere(seed + (real)i))), EPS);
        v3 ca, tn = vs(0); real cm;
        ORvoxel nv;
        memset(&nv, 0, sizeof(nv));
        int hit = step_map(s, c->t, &dir, vinv(dir), &p, 1, &tn, &nv, &ca, &cm);
        if(hit)
        {
            hv = nv;
            hitNormal = hv.normal;
            hm = &s->mats[hv.material];
            v3 dist = vabs(vsub(vfloor(vmuls(lastPos, RL(8.0))), vfloor(vmuls(p, RL(8.0)))));
            if(vdot(dist, dist) < RL(1.0)) return;
            if(hm->emissive)
            {
                *color = vadd(*color, vmul(newColor, vadd(vmuls(vmuls(hv.albedo, emission(hm)), cm), ca)));
                return;
            }
            newColor = vmul(newColor, vadd(vmuls(hv.albedo, cm), ca));
        }
        else
        {
            real cosv = gmax(vdot(dir, c->sunDir), RL(0.0));
            *color = vadd(*color, vadd(vmuls(vmul(vmuls(newColor, cosv), c->sunStrength), cm), ca));
            return;
        }
        lastDir = dir;
    }
}

static inline uint32_t rne_u(real x) { return (uint32_t)r_rint(x); }

/* Runs one lighting dispatch over `requests`. Reads from the scene snapshot,
 * writes lit voxel words into voxelsOut (a full copy of the pool), and applies
 * the pass epilogue to mapOut / chunksOut (copies of map / chunks):
 * visible bit cleared on lit chunks, then set on chunks reached by specular
 * rays from visible chunks; sample count +1 on lit chunks.
 * preq (optional, numRequests*32*6) receives the pre-quantisation diffuse and
 * specular values, voxOf (optional, numRequests*32) the pool index per lane. */
/* Optional output of or_lighting: numRequests*32*3 reals receiving each voxel's
 * diffuse samples of the pass averaged before the running mean
 * (voxelLighting.comp:254-261); lanes without a diffuse estimate are left as
 * they were. NULL disables. */
static real* g_freshOut = NULL;
void or_set_fresh_output(real* out) { g_freshOut = out; }

/* One contiguous range of requests of a lighting dispatch. Reads only the
 * dispatch-start state; writes voxel words, pre-quantisation values and pool
 * indices of its own slots, and its own per-tile arrays (lit, visible
 * propagation, request and lastUsed writes), so ranges are independent. */
typedef struct LRange {
    const ORscene* s;
    const ORlightParams* lp;
    const uint32_t* requests;
    uint32_t r0, r1;
    uint32_t* voxelsOut;
    real* preq;
    uint32_t* voxOf;
    uint8_t* prop;
    uint8_t* lit;
    uint32_t* flagsW;
    uint32_t* lastW;
    const int32_t* colTop;        /* or_present_column_tops of the dispatch-start state */
    ORcounters cnt;
} LRange;

static void light_range(LRange* j)
{
    const ORscene* s = j->s;
    const ORlightParams* lp = j->lp;
    ORscene sc = *s;
    sc.flagsOut = j->flagsW;
    sc.lastUsedOut = j->lastW;
    ORthread t;
    thread_init(&t, 0);
    t.stopOutside = or_stop_outside;
    present_box(s, t.presLo, t.presHi);
    t.colTop = j->colTop;
    LCtx c;
    c.s = &sc; c.lp = lp; c.t = &t; c.visProp = j->prop;
    c.sunDir = vf(lp->sunDir);
    c.sunStrength = vf(s->sunStrength);
    v3 amb = vf(lp->ambient);
    for(uint32_t r = j->r0; r < j->r1; r++)
    {
        uint32_t mapIndex = j->requests[r] >> 4;
        uint32_t block = j->requests[r] & 15u;
        const uint32_t* chunk = s->chunks + (size_t)mapIndex * CHUNK_WORDS;
        j->lit[mapIndex] = 1;
        for(uint32_t lane = 0; lane < 32; lane++)
        {
            size_t slot = (size_t)r * 32 + lane;
            if(j->voxOf) j->voxOf[slot] = 0xFFFFFFFFu;
            uint32_t voxNum = lane + block * 32;
            int32_t sel = or_select(chunk, voxNum);
            if(sel < 0) continue;
            t.lastVoxID = 255u; t.lastVoxRefract = RL(1.0);
            iv3 cp = {sel % 8, (sel / 8) % 8, sel / 64};
            iv3 mp = {(int32_t)chunk[0], (int32_t)chunk[1], (int32_t)chunk[2]};
            uint32_t vi = s->map[(size_t)mapIndex * TILE_WORDS + 2] + (uint32_t)voxNum;
            const uint32_t* cw = s->voxels + (size_t)vi * VOX_WORDS;
            ORvoxel me = decompress(cw);
            const ORmaterial* mat = &s->mats[me.material];
            uint32_t ns = chunk[3];
            real indirect = (real)(ns < lp->maxDiffuseSamples ? ns : lp->maxDiffuseSamples);
            c.firstSample = indirect == 0;
            v3 p = vadd(vadd(vmuls(vfromi(cp), RL(0.125)), vfromi(mp)), vs(RL(0.0625)));
            p = vadd(p, vmul(vs(RL(0.0625) - EPS), me.normal));
            v3 spec = vs(0), diff = vs(0);
            v3 view = vsub(p, vf(lp->camPos));
            if(mat->specular > 0.0f && vdot(view, me.normal) < 0 && mat->reflectType <= 1u)
            {
                v3 refl = vreflect(vnorm(view), me.normal);
                for(uint32_t i = 0; i < lp->numSpecRays; i++)
                {
                    v3 sd = vadds(vnorm(vadd(vmuls(refl, (real)mat->shininess), vf(lp->specPoints + 3 * i))), EPS);
                    specular_ray(&c, mapIndex, p, sd, me.albedo, mat->reflectType, &spec);
                }
                spec = vdivs(spec, (real)lp->numSpecRays);
            }
            if(mat->specular < 1.0f)
            {
                for(uint32_t i = 0; i < lp->numDiffuseSamples; i++)
                {
                    diff = vadd(diff, amb);
                    diffuse_ray(&c, vadds(me.normal, EPS), p, me, (real)lp->time * (real)(i + 1), &diff);
                    shadow_ray(&c, p, (real)lp->time * (real)(i + 1 + lp->numDiffuseSamples), &diff);
                }
                if(g_freshOut && lp->numDiffuseSamples)
                {
                    v3 fr = vdivs(diff, (real)lp->numDiffuseSamples);
                    g_freshOut[slot * 3 + 0] = fr.x; g_freshOut[slot * 3 + 1] = fr.y; g_freshOut[slot * 3 + 2] = fr.z;
                }
                diff = vdivs(vadd(vmuls(me.diffuse, indirect), diff), indirect + (real)lp->numDiffuseSamples);
            }
            if(j->preq)
            {
                real* q = j->preq + slot * 6;
                q[0] = diff.x; q[1] = diff.y; q[2] = diff.z;
                q[3] = spec.x; q[4] = spec.y; q[5] = spec.z;
            }
            if(j->voxOf) j->voxOf[slot] = vi;
            spec = vclamp(spec, RL(0.0), RL(1.0));
            diff = vclamp(diff, RL(0.0), RL(1.0));
            uint32_t wdx = rne_u(diff.x * RL(65535.0)), wdy = rne_u(diff.y * RL(65535.0)), wdz = rne_u(diff.z * RL(65535.0));
            uint32_t ar = rne_u(me.albedo.x * RL(255.0)) & 255u, ag = rne_u(me.albedo.y * RL(255.0)) & 255u,
                     ab = rne_u(me.albedo.z * RL(255.0)) & 255u;
            uint32_t sx = rne_u(spec.x * RL(255.0)) & 255u, sy = rne_u(spec.y * RL(255.0)) & 255u,
                     sz = rne_u(spec.z * RL(255.0)) & 255u;
            uint32_t* ow = j->voxelsOut + (size_t)vi * VOX_WORDS;
            ow[1] = (ar << 24) | (ag << 16) | (ab << 8) | sx;
            ow[2] = (sy << 24) | (sz << 16) | (((wdx >> 8) & 255u) << 8) | (wdx & 255u);
            ow[3] = (((wdy >> 8) & 255u) << 24) | ((wdy & 255u) << 16) | (((wdz >> 8) & 255u) << 8) | (wdz & 255u);
        }
    }
    j->cnt = t.cnt;
}

static void* light_range_thread(void* arg)
{
    light_range((LRange*)arg);
    return NULL;
}

/* The lighting dispatch (voxelLighting.comp:207-288) over `numRequests`
 * requests, split into `numThreads` contiguous ranges run on host threads.
 * Every range reads the dispatch-start state only, so the result does not
 * depend on the split: the per-tile writes are merged afterwards (a lit tile
 * loses its visible bit, a tile reached by specular rays gains it, a request
 * write sets flags 3, a loaded tile entered by a ray gets lastUsed 0). */
void or_lighting_mt(const ORscene* s, const ORlightParams* lp, const uint32_t* requests, uint32_t numRequests,
                    uint32_t* voxelsOut, uint32_t* mapOut, uint32_t* chunksOut, real* preq, uint32_t* voxOf,
                    ORcounters* cnt, uint32_t numThreads)
{
    uint32_t tiles = s->mapSize[0] * s->mapSize[1] * s->mapSize[2];
    size_t tw = tiles ? tiles : 1;
    if(numThreads < 1) numThreads = 1;
    if(numThreads > numRequests) numThreads = numRequests ? numRequests : 1;
    LRange* jobs = (LRange*)calloc(numThreads, sizeof(LRange));
    const size_t plane = (size_t)s->mapSize[0] * s->mapSize[2];
    int32_t* colTop = (int32_t*)malloc((plane > 0 ? 4 * plane : 1) * sizeof(int32_t));
    or_present_column_tops(s, colTop);
    for(uint32_t k = 0; k < numThreads; k++)
    {
        LRange* j = &jobs[k];
        j->s = s; j->lp = lp; j->requests = requests; j->colTop = colTop;
        j->r0 = (uint32_t)((uint64_t)numRequests * k / numThreads);
        j->r1 = (uint32_t)((uint64_t)numRequests * (k + 1) / numThreads);
        j->voxelsOut = voxelsOut; j->preq = preq; j->voxOf = voxOf;
        j->prop = (uint8_t*)calloc(tw, 1);
        j->lit = (uint8_t*)calloc(tw, 1);
        j->flagsW = (uint32_t*)malloc(sizeof(uint32_t) * tw);
        j->lastW = (uint32_t*)malloc(sizeof(uint32_t) * tw);
        for(uint32_t i = 0; i < tiles; i++) { j->flagsW[i] = s->map[3 * i]; j->lastW[i] = s->map[3 * i + 1]; }
    }
    if(numThreads == 1) light_range(&jobs[0]);
    else
    {
        pthread_t* th = (pthread_t*)calloc(numThreads, sizeof(pthread_t));
        for(uint32_t k = 0; k < numThreads; k++) pthread_create(&th[k], NULL, light_range_thread, &jobs[k]);
        for(uint32_t k = 0; k < numThreads; k++) pthread_join(th[k], NULL);
        free(th);
    }
    ORcounters total = {0, 0, 0, 0};
    for(uint32_t k = 0; k < numThreads; k++)
    {
        total.calls += jobs[k].cnt.calls; total.sMap += jobs[k].cnt.sMap;
        total.sVox += jobs[k].cnt.sVox; total.hits += jobs[k].cnt.hits;
    }
    for(uint32_t i = 0; i < tiles; i++)
    {
        uint32_t f = s->map[3 * i], last = s->map[3 * i + 1];
        int lit = 0, prop = 0;
        for(uint32_t k = 0; k < numThreads; k++)
        {
            const LRange* j = &jobs[k];
            if(j->flagsW[i] != s->map[3 * i]) f = j->flagsW[i];
            if(j->lastW[i] < last) last = j->lastW[i];
            lit |= j->lit[i];
            prop |= j->prop[i];
        }
        if(lit) f &= ~4u;
        if(prop) f |= 4u;
        mapOut[3 * i] = f;
        mapOut[3 * i + 1] = last;
        mapOut[3 * i + 2] = s->map[3 * i + 2];
        if(lit) chunksOut[(size_t)i * CHUNK_WORDS + 3] = s->chunks[(size_t)i * CHUNK_WORDS + 3] + 1u;
    }
    for(uint32_t k = 0; k < numThreads; k++)
    {
        free(jobs[k].prop); free(jobs[k].lit); free(jobs[k].flagsW); free(jobs[k].lastW);
    }
    free(jobs);
    free(colTop);
    if(cnt) *cnt = total;
}

void or_lighting(const ORscene* s, const ORlightParams* lp, const uint32_t* requests, uint32_t numRequests,
                 uint32_t* voxelsOut, uint32_t* mapOut, uint32_t* chunksOut, real* preq, uint32_t* voxOf,
                 ORcounters* cnt)
{
    or_lighting_mt(s, lp, requests, numRequests, voxelsOut, mapOut, chunksOut, preq, voxOf, cnt, 1);
}

/* ------------------------------------------------------------------------- */
/* chunk conversion (voxel.c:1391-1461). Host C code in the reference: always  */
/* FP32, whatever the oracle's `real` type is.                                 */

/* cpuVoxels: DNchunk.voxels[8][8][8] of {normal, albedo}, flat index ((x*8)+y)*8+z.
 * Writes chunkOut (24 words; pos and sample count left to the caller) and up to
 * 512 voxels (4 words each); returns the number of stored voxels. */
uint32_t or_chunk_to_gpu(const ORmaterial* mats, const uint32_t* cpuVoxels, uint32_t* chunkOut, uint32_t* voxOut)
{
    for(int i = 0; i < 16; i++) chunkOut[7 + i] = 0;
    uint32_t n = 0;
#define CV(x, y, z) (cpuVoxels + 2 * ((((x) * 8) + (y)) * 8 + (z)))
#define FACE(x, y, z) (!((x) >= 0 && (x) < 8 && (y) >= 0 && (y) < 8 && (z) >= 0 && (z) < 8) || \
                       (CV(x, y, z)[0] >> 24) == 255u || mats[CV(x, y, z)[0] >> 24].opacity < 1.0f)
    for(int z = 0; z < 8; z++)
    for(int y = 0; y < 8; y++)
    for(int x = 0; x < 8; x++)
    {
        uint32_t index = (uint32_t)(x + 8 * (y + 8 * z));
        if((index & 31u) == 0 && index != 0 && ((index >> 5) & 3u) == 0) chunkOut[4 + (index >> 7) - 1] = n;
        const uint32_t* v = CV(x, y, z);
        if((v[0] >> 24) == 255u) continue;
        int vis = FACE(x + 1, y, z) || FACE(x - 1, y, z) || FACE(x, y + 1, z) || FACE(x, y - 1, z) ||
                  FACE(x, y, z + 1) || FACE(x, y, z - 1);
        if(!vis) continue;
        chunkOut[7 + (index >> 5)] |= 1u << (index & 31u);
        uint32_t a = v[1];
        float lx = (float)((a >> 24) & 255u) * 0.00392156862f;
        float ly = (float)((a >> 16) & 255u) * 0.00392156862f;
        float lz = (float)((a >> 8) & 255u) * 0.00392156862f;
        lx = powf(lx, 2.2f) * 255.0f; ly = powf(ly, 2.2f) * 255.0f; lz = powf(lz, 2.2f) * 255.0f;
        uint8_t r = (uint8_t)lx, g = (uint8_t)ly, b = (uint8_t)lz;
        uint32_t* o = voxOut + 4 * n;
        o[0] = v[0];
        o[1] = ((uint32_t)r << 24) | ((uint32_t)g << 16) | ((uint32_t)b << 8);
        o[2] = 0; o[3] = 0;
        n++;
    }
#undef FACE
#undef CV
    return n;
}

/* ------------------------------------------------------------------------- */
/* visible-chunk request list (voxel.c:719-762, 1463-1489)                     */

/* Tiles in z,y,x order with loaded flag 2 and the visible bit, that are in
 * this frame's lighting split or were (re)uploaded this sync (fresh != 0),
 * emit one word (mapIndex << 4) | block per 32 stored voxels. The stored
 * voxel count comes from the chunk's occupancy mask. */
uint32_t or_visible_requests(const uint32_t* map, const uint32_t* chunks, uint32_t tiles, const uint8_t* fresh,
                             uint32_t split, uint32_t frameNum, uint32_t* out)
{
    uint32_t k = 0;
    for(uint32_t i = 0; i < tiles; i++)
    {
        uint32_t f = map[3 * i];
        if((f & 3u) != 2u || !(f & 4u)) continue;
        if(i % split != frameNum && !(fresh && fresh[i])) continue;
        uint32_t n = 0;
        for(int w = 0; w < 16; w++) n += popc(chunks[(size_t)i * CHUNK_WORDS + 7 + w]);
        for(uint32_t b = 0; b * 32 < n; b++) out[k++] = (i << 4) | b;
    }
    return k;
}

/* sizes of the exchanged structs, for the ctypes mirror */
uint32_t or_sizeof(int which)
{
    switch(which)
    {
    case 0: return sizeof(ORmaterial);
    case 1: return sizeof(ORscene);
    case 2: return sizeof(ORcounters);
    case 3: return sizeof(ORrayHit);
    case 4: return sizeof(ORdrawParams);
    case 5: return sizeof(ORpixel);
    case 6: return sizeof(ORlightParams);
    case 7: return sizeof(real);
    default: return 0;
    }
}

/* decompress_voxel for tests: out = normal(3), albedo(3), spec(3), diffuse(3), material */
void or_decompress(const uint32_t* w, real* out)
{
    ORvoxel v = decompress(w);
    out[0] = v.normal.x; out[1] = v.normal.y; out[2] = v.normal.z;
    out[3] = v.albedo.x; out[4] = v.albedo.y; out[5] = v.albedo.z;
    out[6] = v.spec.x; out[7] = v.spec.y; out[8] = v.spec.z;
    out[9] = v.diffuse.x; out[10] = v.diffuse.y; out[11] = v.diffuse.z;
    out[12] = (real)v.material;
}

/* ---------------------------------------------------------------------------
 * Host voxel edits on a dense grid (always FP32, as the reference's C code).
 * The grid holds X*Y*Z voxels (X = 8 * map width ...), voxel (x, y, z) at
 * x + X*(y + Y*z); per tile a chunk-present flag and its stored-voxel count.
 * A tile without a chunk holds the cleared words (normal 0xFFFFFFFF, albedo 0),
 * the state _DN_clear_chunk leaves (voxel.c:1353-1363), which is also what a
 * newly added chunk starts from. Chunk pool slots are not modelled.
 */
typedef struct {
    int32_t mx, my, mz;        /* map size in tiles */
    uint32_t* normal;          /* per voxel */
    uint32_t* albedo;          /* per voxel */
    uint8_t* flag;             /* per tile: chunk present */
    int32_t* count;            /* per tile: stored voxels */
} ORgrid;

static inline uint32_t g_vox(const ORgrid* g, int x, int y, int z)
{
    return (uint32_t)x + (uint32_t)(8 * g->mx) * ((uint32_t)y + (uint32_t)(8 * g->my) * (uint32_t)z);
}

static void g_clear_tile(ORgrid* g, int tx, int ty, int tz)
{
    for(int z = 0; z < 8; z++)
    for(int y = 0; y < 8; y++)
    for(int x = 0; x < 8; x++)
    {
        uint32_t v = g_vox(g, tx * 8 + x, ty * 8 + y, tz * 8 + z);
        g->normal[v] = 0xFFFFFFFFu;
        g->albedo[v] = 0u;
    }
}

/* DN_set_compressed_voxel, voxel.c:1126-1160: an empty voxel into a missing
 * chunk does nothing, a solid one adds the chunk; the count follows empty<->solid
 * changes and the chunk is removed (cleared) when it reaches zero. */
void or_set_voxel_word(ORgrid* g, int x, int y, int z, uint32_t normal, uint32_t albedo)
{
    int tx = x / 8, ty = y / 8, tz = z / 8;
    uint32_t t = (uint32_t)tx + (uint32_t)g->mx * ((uint32_t)ty + (uint32_t)g->my * (uint32_t)tz);
    uint32_t v = g_vox(g, x, y, z);
    if(!g->flag[t])
    {
        if((normal >> 24) == 255u) return;
        g->flag[t] = 1;
    }
    uint32_t oldMat = g->normal[v] >> 24, newMat = normal >> 24;
    if(oldMat == 255u && newMat != 255u) g->count[t]++;
    else if(oldMat != 255u && newMat == 255u)
    {
        if(--g->count[t] <= 0)
        {
            g->flag[t] = 0;
            g->count[t] = 0;
            g_clear_tile(g, tx, ty, tz);
            return;
        }
    }
    g->normal[v] = normal;
    g->albedo[v] = albedo;
}

/* DNX_set_voxels_dense: or_set_voxel_word over a box of sx*sy*sz voxels from
 * (ox, oy, oz), x fastest, then y, then z (the words in that order). */
void or_set_dense(ORgrid* g, int ox, int oy, int oz, int sx, int sy, int sz, const uint32_t* normal,
                  const uint32_t* albedo)
{
    size_t k = 0;
    for(int z = 0; z < sz; z++)
    for(int y = 0; y < sy; y++)
    for(int x = 0; x < sx; x++, k++)
        or_set_voxel_word(g, ox + x, oy + y, oz + z, normal[k], albedo[k]);
}

/* DN_compress_voxel, voxel.c:1289-1301 */
static uint32_t g_compress_normal(uint32_t material, float nx, float ny, float nz)
{
    nx = nx < 1.0f ? nx : 1.0f; ny = ny < 1.0f ? ny : 1.0f; nz = nz < 1.0f ? nz : 1.0f;
    nx = nx > -1.0f ? nx : -1.0f; ny = ny > -1.0f ? ny : -1.0f; nz = nz > -1.0f ? nz : -1.0f;
    uint32_t qx = (uint32_t)(((int)(nx * 255.0f) + 255) / 2), qy = (uint32_t)(((int)(ny * 255.0f) + 255) / 2),
             qz = (uint32_t)(((int)(nz * 255.0f) + 255) / 2);
    return (material << 24) | (qx << 16) | (qy << 8) | qz;
}

/* The transform callback's view of a voxel and vectors: DNvoxel / DNvec3 /
 * DNmat4 of voxel.h and quickmath.h, by value (voxelShapes.h:28). */
typedef struct { float x, y, z; } ORv3f;
typedef struct { uint8_t material; float normal[3]; uint8_t albedo[3]; } ORvoxelT;
typedef struct { float m[4][4]; } ORmat4;
typedef ORvoxelT (*ORtransformFn)(ORv3f pos, ORv3f normal, ORvoxelT vox, ORv3f shapeMin, ORv3f shapeMax,
                                  ORmat4 invTransform, void* userData);

/* voxelShapes.c:242-246: inverse of translate(c) * quaternion_to_mat4(orient),
 * QuickMath's translate (quickmath.h:1138-1147), quaternion conversion
 * (:1622-1650), column-major product (:884-917) and cofactor inverse
 * (:1037-1124), column-major m[col][row]. */
static void sphere_inv_transform(float cx, float cy, float cz, const float* q, float out[4][4])
{
    float R[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    if(q)
    {
        float x2 = q[0] + q[0], y2 = q[1] + q[1], z2 = q[2] + q[2];
        float xx2 = q[0] * x2, xy2 = q[0] * y2, xz2 = q[0] * z2, yy2 = q[1] * y2, yz2 = q[1] * z2, zz2 = q[2] * z2;
        float sx2 = q[3] * x2, sy2 = q[3] * y2, sz2 = q[3] * z2;
        R[0][0] = 1.0f - (yy2 + zz2); R[0][1] = xy2 - sz2; R[0][2] = xz2 + sy2;
        R[1][0] = xy2 + sz2; R[1][1] = 1.0f - (xx2 + zz2); R[1][2] = yz2 - sx2;
        R[2][0] = xz2 - sy2; R[2][1] = yz2 + sx2; R[2][2] = 1.0f - (xx2 + yy2);
    }
    float T[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {cx, cy, cz, 1}};
    float M[4][4];
    for(int c = 0; c < 4; c++)
        for(int r = 0; r < 4; r++)
            M[c][r] = T[0][r] * R[c][0] + T[1][r] * R[c][1] + T[2][r] * R[c][2] + T[3][r] * R[c][3];
    float a = M[0][0], b = M[0][1], c = M[0][2], d = M[0][3], e = M[1][0], f = M[1][1], g = M[1][2], h = M[1][3],
          i = M[2][0], j = M[2][1], k = M[2][2], l = M[2][3], m = M[3][0], n = M[3][1], o = M[3][2], p = M[3][3];
    float t0 = k * p - o * l, t1 = j * p - n * l, t2 = j * o - n * k, t3 = i * p - m * l, t4 = i * o - m * k,
          t5 = i * n - m * j;
    out[0][0] = f * t0 - g * t1 + h * t2;
    out[1][0] = -(e * t0 - g * t3 + h * t4);
    out[2][0] = e * t1 - f * t3 + h * t5;
    out[3][0] = -(e * t2 - f * t4 + g * t5);
    out[0][1] = -(b * t0 - c * t1 + d * t2);
    out[1][1] = a * t0 - c * t3 + d * t4;
    out[2][1] = -(a * t1 - b * t3 + d * t5);
    out[3][1] = a * t2 - b * t4 + c * t5;
    t0 = g * p - o * h; t1 = f * p - n * h; t2 = f * o - n * g; t3 = e * p - m * h; t4 = e * o - m * g; t5 = e * n - m * f;
    out[0][2] = b * t0 - c * t1 + d * t2;
    out[1][2] = -(a * t0 - c * t3 + d * t4);
    out[2][2] = a * t1 - b * t3 + d * t5;
    out[3][2] = -(a * t2 - b * t4 + c * t5);
    t0 = g * l - k * h; t1 = f * l - j * h; t2 = f * k - j * g; t3 = e * l - i * h; t4 = e * k - i * g; t5 = e * j - i * f;
    out[0][3] = -(b * t0 - c * t1 + d * t2);
    out[1][3] = a * t0 - c * t3 + d * t4;
    out[2][3] = -(a * t1 - b * t3 + d * t5);
    out[3][3] = a * t2 - b * t4 + c * t5;
    float det = 1.0f / (a * out[0][0] + b * out[1][0] + c * out[2][0] + d * out[3][0]);
    for(int x = 0; x < 4; x++)
        for(int y = 0; y < 4; y++) out[x][y] = out[x][y] * det;
}

/* DN_shape_sphere, voxelShapes.c:239-333: every
 * voxel of the (clamped) tile box, tile by tile in z, y, x order and voxel by
 * voxel in z, y, x order inside a tile; FP32 squared distance from the centre,
 * dot = (x*x + y*y) + z*z. Solid spheres skip voxels that exist and write the
 * normal fromCenter / max|fromCenter|; empty ones write the caller's (empty)
 * voxel over every voxel inside and,
 * with flipNormals, turn the normals of the voxels in the next shell inwards.
 * The chunk-present flag is read once per tile, as the reference does. */
void or_shape_sphere_fn(ORgrid* g, uint32_t material, const float* normal, uint32_t albedoWord, int flipNormals,
                        float cx, float cy, float cz, float r, const float* orient, ORtransformFn fn, void* userData)
{
    ORmat4 inv;
    sphere_inv_transform(cx, cy, cz, orient, inv.m);
    const ORv3f shapeMin = {cx - r, cy - r, cz - r}, shapeMax = {cx + r, cy + r, cz + r};
    int iMin[3] = {(int)floorf(cx - r), (int)floorf(cy - r), (int)floorf(cz - r)};
    int iMax[3] = {(int)ceilf(cx + r), (int)ceilf(cy + r), (int)ceilf(cz + r)};
    if(material == 255u)
        for(int a = 0; a < 3; a++) { iMin[a]--; iMax[a]++; }
    int size[3] = {g->mx, g->my, g->mz}, lo[3], hi[3];
    for(int a = 0; a < 3; a++)
    {
        lo[a] = iMin[a] / 8;              /* C division: truncation, as DN_separate_position */
        hi[a] = iMax[a] / 8;
        if(lo[a] < 0) lo[a] = 0;
        if(hi[a] > size[a] - 1) hi[a] = size[a] - 1;
    }
    const float r2 = r * r, r12 = (r + 1.0f) * (r + 1.0f);
    for(int mZ = lo[2]; mZ <= hi[2]; mZ++)
    for(int mY = lo[1]; mY <= hi[1]; mY++)
    for(int mX = lo[0]; mX <= hi[0]; mX++)
    {
        uint32_t t = (uint32_t)mX + (uint32_t)g->mx * ((uint32_t)mY + (uint32_t)g->my * (uint32_t)mZ);
        int tileFlag = g->flag[t];
        for(int cZ = 0; cZ < 8; cZ++)
        for(int cY = 0; cY < 8; cY++)
        for(int cX = 0; cX < 8; cX++)
        {
            int x = mX * 8 + cX, y = mY * 8 + cY, z = mZ * 8 + cZ;
            float fx = (float)x - cx, fy = (float)y - cy, fz = (float)z - cz;
            float d2 = fx * fx + fy * fy + fz * fz;
            uint32_t v = g_vox(g, x, y, z);
            if(d2 < r2)
            {
                if(tileFlag && (g->normal[v] >> 24) != 255u && material != 255u) continue;
                float nx = normal[0], ny = normal[1], nz = normal[2];
                uint32_t mt = material, aw = albedoWord;
                if(material != 255u)
                {
                    float m = fmaxf(fmaxf(fabsf(fx), fabsf(fy)), fabsf(fz));
                    float s = 1.0f / m;
                    nx = fx * s; ny = fy * s; nz = fz * s;
                    if(fn)
                    {
                        ORvoxelT in = {(uint8_t)material, {normal[0], normal[1], normal[2]},
                                       {(uint8_t)(albedoWord >> 24), (uint8_t)(albedoWord >> 16), (uint8_t)(albedoWord >> 8)}};
                        ORv3f pos = {(float)x, (float)y, (float)z}, n = {nx, ny, nz};
                        ORvoxelT o = fn(pos, n, in, shapeMin, shapeMax, inv, userData);
                        mt = o.material; nx = o.normal[0]; ny = o.normal[1]; nz = o.normal[2];
                        aw = ((uint32_t)o.albedo[0] << 24) | ((uint32_t)o.albedo[1] << 16) | ((uint32_t)o.albedo[2] << 8);
                    }
                }
                or_set_voxel_word(g, x, y, z, g_compress_normal(mt, nx, ny, nz), aw);
            }
            else if(material == 255u && flipNormals && d2 < r12)
            {
                if(!tileFlag || (g->normal[v] >> 24) == 255u) continue;
                float m = fmaxf(fmaxf(fabsf(fx), fabsf(fy)), fabsf(fz));
                float s = -1.0f / m;
                if(fn)
                {
                    /* DN_get_voxel decodes the old voxel (voxel.c:1304-1316), func gets it */
                    const float inv255 = 1.0f / 255.0f;
                    uint32_t w = g->normal[v], al = g->albedo[v];
                    ORvoxelT old = {(uint8_t)(w >> 24),
                                    {(float)((int)((w >> 16) & 255u) * 2 - 255) * inv255,
                                     (float)((int)((w >> 8) & 255u) * 2 - 255) * inv255,
                                     (float)((int)(w & 255u) * 2 - 255) * inv255},
                                    {(uint8_t)(al >> 24), (uint8_t)(al >> 16), (uint8_t)(al >> 8)}};
                    ORv3f pos = {(float)x, (float)y, (float)z}, n = {fx * s, fy * s, fz * s};
                    ORvoxelT o = fn(pos, n, old, shapeMin, shapeMax, inv, userData);
                    or_set_voxel_word(g, x, y, z, g_compress_normal(o.material, o.normal[0], o.normal[1], o.normal[2]),
                                      ((uint32_t)o.albedo[0] << 24) | ((uint32_t)o.albedo[1] << 16) |
                                          ((uint32_t)o.albedo[2] << 8));
                    continue;
                }
                /* DN_get_voxel then DN_set_voxel: the albedo keeps its three colour bytes */
                or_set_voxel_word(g, x, y, z, g_compress_normal(g->normal[v] >> 24, fx * s, fy * s, fz * s),
                                  g->albedo[v] & 0xFFFFFF00u);
            }
        }
    }
}

void or_shape_sphere(ORgrid* g, uint32_t material, const float* normal, uint32_t albedoWord, int flipNormals,
                     float cx, float cy, float cz, float r)
{
    or_shape_sphere_fn(g, material, normal, albedoWord, flipNormals, cx, cy, cz, r, NULL, NULL, NULL);
}

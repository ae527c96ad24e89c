// Learned radiance cache on gfx950 (no counterpart in the reference).
//
// Encoding: multiresolution hash grid (Mueller et al., "Instant Neural Graphics
// Primitives with a Multiresolution Hash Encoding", 2022, sec. 3): per level l
// the 8 corners of the cell around x * N_l + 0.5, trilinear weights, a dense
// index where (N_l + 2)^3 entries fit the table, else the spatial hash
// (x * 1 ^ y * 2654435761 ^ z * 805459861) mod T.
// Network: 4 weight matrices without biases, ReLU between them; operands in
// bfloat16 on v_mfma_f32_16x16x32_bf16 (v_mfma_f32_16x16x16_bf16 for a
// lone 16-wide K block and for the weight-gradient GEMM) with FP32
// accumulation. Layers are
// computed transposed (D[feature][sample] = W * X^T), so one layer's 16x16
// accumulator tile is, lane for lane, the next layer's B operand: a lane holds
// sample (lane & 15) and features 4 * (lane >> 4) + 0..3 of a 16-feature block.
// Training: forward + backward per 16-sample tile, weight gradients as a
// split-K MFMA reduction over the tiles, hash-table gradients by FP32 atomics,
// then Adam (Kingma & Ba 2015, algorithm 1) on FP32 master copies.
#include "dn_light.hpp"
#include "rc_internal.h"
#include "doon/doon.h"

#include <mutex>
#include <vector>

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two consecutive K = 16 steps as one v_mfma_f32_16x16x32_bf16 (gfx950):
// c + A0 B0 + A1 B1. A lane group's 8 k-values of A are its 4 of A0 then its
// 4 of A1, and likewise for B, so A and B pair the same 32 k-values (a
// permutation of the K = 32 block, which leaves the dot product a sum over
// the same terms) and two C-layout tiles feed B without any shuffle.
__device__ __forceinline__ f32x4 mfma32(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 c)
{
    const s16x8 a = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
    const s16x8 b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

__device__ __forceinline__ bf16x4 ld_bf16x4(const uint16_t* p)
{
    uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFFu); r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFFu); r[3] = (short)(v.y >> 16);
    return r;
}

// A operand of a 16x16 block of a row-major bf16 matrix M (leading dimension
// ld) at (row0, k0): lane holds M[row0 + (lane & 15)][k0 + 4 * (lane >> 4) + j].
__device__ __forceinline__ bf16x4 a_rows(const uint16_t* M, uint32_t ld, uint32_t row0, uint32_t k0, uint32_t lane)
{
    return ld_bf16x4(M + (row0 + (lane & 15u)) * ld + k0 + 4u * (lane >> 4));
}

__device__ __forceinline__ bf16x4 relu_bf16(f32x4 c)
{
    bf16x4 r;
    for(int i = 0; i < 4; i++) r[i] = (short)f2bf(c[i] > 0.0f ? c[i] : 0.0f);
    return r;
}

// Gradient through ReLU: keep d where the layer output h (bf16) is positive.
__device__ __forceinline__ bf16x4 relu_grad(f32x4 d, bf16x4 h)
{
    bf16x4 r;
    for(int i = 0; i < 4; i++)
    {
        uint16_t hb = (uint16_t)h[i];
        r[i] = (short)((hb != 0u && !(hb & 0x8000u)) ? f2bf(d[i]) : 0u);
    }
    return r;
}

// Stores a C-layout bf16 tile of feature block fb into a tile-transposed
// buffer T[tile][F][16].
__device__ __forceinline__ void store_T(uint16_t* T, uint32_t F, uint32_t tile, uint32_t fb, bf16x4 v, uint32_t lane)
{
    uint16_t* p = T + ((size_t)tile * F + fb * 16u + 4u * (lane >> 4)) * 16u + (lane & 15u);
    p[0] = (uint16_t)v[0];
    p[16] = (uint16_t)v[1];
    p[32] = (uint16_t)v[2];
    p[48] = (uint16_t)v[3];
}

// ---------------------------------------------------------------------------
// encoding

// The cell of a level that holds a position: its lower corner g and the
// position's fractions f inside it.
struct LevelCell {
    uint32_t g[3];
    float f[3];
};

__device__ __forceinline__ LevelCell level_cell(const DNrcNet& net, uint32_t l, const float* pos)
{
    LevelCell c;
    const uint32_t N = net.res[l];
    for(int a = 0; a < 3; a++)
    {
        float x = pos[a] / net.extent;
        x = fminf(fmaxf(x, 0.0f), 1.0f);
        float s = x * (float)N + 0.5f;
        float fl = floorf(s);
        c.g[a] = (uint32_t)fl;
        c.f[a] = s - fl;
    }
    return c;
}

// Corner k (bit 0: x + 1, bit 1: y + 1, bit 2: z + 1) of a cell: its table
// entry (dense index, or the spatial hash) and its trilinear weight.
__device__ __forceinline__ void cell_corner(const DNrcNet& net, uint32_t l, const LevelCell& c, uint32_t k,
                                            uint32_t& idx, float& w)
{
    const uint32_t T = 1u << net.log2T;
    const bool dense = (net.denseMask >> l) & 1u;
    const uint32_t S = net.res[l] + 2u;
    const uint32_t cx = c.g[0] + (k & 1u), cy = c.g[1] + ((k >> 1) & 1u), cz = c.g[2] + ((k >> 2) & 1u);
    const uint32_t i = dense ? cx + cy * S + cz * S * S : (cx * 1u) ^ (cy * 2654435761u) ^ (cz * 805459861u);
    idx = i & (T - 1u);
    w = ((k & 1u) ? c.f[0] : 1.0f - c.f[0]) * ((k & 2u) ? c.f[1] : 1.0f - c.f[1]) *
        ((k & 4u) ? c.f[2] : 1.0f - c.f[2]);
}

// One thread per (sample, level): the level's two features into x0, level 0
// also the normal, albedo and zero padding. (Two lanes per pair, each
// gathering the corners of one x offset so that a load instruction takes a
// sample's two x-neighbour entries at once, measured 52.7 against 46.9 us per
// 2^17 samples: unlike the atomics, these gathers are not bound by the cache
// lines an instruction touches.)
__global__ __launch_bounds__(256) void k_rc_encode(DNrcNet net, const float* samples, uint32_t n, uint16_t* x0)
{
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    const uint32_t s = g >> 4, l = g & 15u;
    if(s >= n) return;
    const float* sp = samples + (size_t)s * RC_SAMPLE;
    const LevelCell cell = level_cell(net, l, sp);
    const uint32_t* tab = net.tab16 + ((size_t)l << net.log2T);
    uint32_t idx[8];
    float w[8];
    for(uint32_t k = 0; k < 8u; k++) cell_corner(net, l, cell, k, idx[k], w[k]);
    uint32_t e[8];
    for(uint32_t k = 0; k < 8u; k++) e[k] = tab[idx[k]];        // all eight gathers in flight at once
    float a0 = 0.0f, a1 = 0.0f;
    for(uint32_t k = 0; k < 8u; k++)
    {
        a0 += w[k] * bf2f(e[k] & 0xFFFFu);
        a1 += w[k] * bf2f(e[k] >> 16);
    }
    uint16_t* row = x0 + (size_t)s * RC_IN;
    row[2 * l] = f2bf(a0);
    row[2 * l + 1] = f2bf(a1);
    if(l == 0)
    {
        for(int i = 0; i < 6; i++) row[RC_ENC + i] = f2bf(sp[3 + i]);
        for(int i = RC_ENC + 6; i < RC_IN; i++) row[i] = 0;
    }
}

// Hash-table gradient, level-major: the waves running at a time work on one
// level's gradient table (4 MiB at 2^19 entries), finest level first.
// Device-scope FP32 atomics are served per cache line that one instruction
// touches, not per lane (measured on MI355X: 64 lanes on 64 scattered lines
// sustain 20.7 G atomics/s, lane pairs on one 8-byte entry 41 G/s, 16 lanes on
// one 64-byte run about 290 G/s). So four lanes take one sample: lane
// q = 2 xb + f adds feature f at the four corners whose x offset is xb. A
// sample's two x-neighbour corners are adjacent entries on a dense level
// (x + S (y + S z)) and, on a hashed level (x ^ h(y, z)), lie in one aligned
// run of entries unless x + 1 carries past it, so each instruction sends two
// corners and both features of a sample to one line. Per 2^17 training
// samples this takes 185 us, against 378 us with one lane per sample (two
// instructions per corner) and per-wave sums of shared entries on the 8
// coarsest levels; with the four-lane layout such sums only cost time
// (8 levels 275 us, 6 levels 239, 4 levels 223).
__global__ __launch_bounds__(256) void k_rc_table_grad(DNrcNet net, const float* samples, uint32_t n, const float* denc)
{
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, w = g >> 6;
    const uint32_t waves = (n + 15u) / 16u;                      // per level
    if(w >= RC_LEVELS * waves) return;
    const uint32_t l = RC_LEVELS - 1u - w / waves;
    const uint32_t s = (w % waves) * 16u + (lane >> 2);
    const uint32_t f = lane & 1u, xb = (lane >> 1) & 1u;
    if(s >= n) return;
    const float d = denc[(size_t)s * RC_ENC + 2u * l + f];
    if(d == 0.0f) return;                                       // adding +-0 changes no sum
    float* gt = net.tabGrad + ((size_t)l << net.log2T) * 2u;
    const LevelCell cell = level_cell(net, l, samples + (size_t)s * RC_SAMPLE);
    for(uint32_t j = 0; j < 4u; j++)
    {
        uint32_t idx;
        float wt;
        cell_corner(net, l, cell, (j << 1) | xb, idx, wt);
        atomicAdd(gt + 2u * idx + f, wt * d);
    }
}

// ---------------------------------------------------------------------------
// MLP: forward (and, with TRAIN, backward) over 16-sample tiles, one tile per
// wave at a time. Weights (and with TRAIN their transposes) staged in LDS.

template <bool TRAIN>
__global__ __launch_bounds__(256) void k_rc_mlp(DNrcNet net, const uint16_t* x0, const float* samples, uint32_t n,
                                                float invB, float* pred, DNrcActs acts)
{
    __shared__ __align__(16) uint16_t W[RC_NPARAM];
    __shared__ __align__(16) uint16_t WT[TRAIN ? RC_NPARAM : 8];
    // the weights, and for the backward pass their transposes (W0^T 48x64,
    // W1^T, W2^T 64x64, W3^T 64x16, same offsets; written by the Adam step),
    // staged in 16-byte pieces
    for(uint32_t i = threadIdx.x; i < RC_NPARAM / 8u; i += 256u)
    {
        reinterpret_cast<uint4*>(W)[i] = reinterpret_cast<const uint4*>(net.w16)[i];
        if(TRAIN) reinterpret_cast<uint4*>(WT)[i] = reinterpret_cast<const uint4*>(net.wT16)[i];
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tiles = n / 16u;
    for(uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6); tile < tiles; tile += gridDim.x * 4u)
    {
        const uint32_t s0 = tile * 16u;
        const uint32_t sm = s0 + (lane & 15u);
        bf16x4 xin[3];
        for(uint32_t ks = 0; ks < 3; ks++) xin[ks] = ld_bf16x4(x0 + (size_t)sm * RC_IN + ks * 16u + 4u * (lane >> 4));
        bf16x4 h[3][4];
        f32x4 acc[4];
        for(uint32_t mb = 0; mb < 4; mb++)
        {
            acc[mb] = mfma32(a_rows(W + RC_OFF_W0, 48, mb * 16u, 0u, lane), a_rows(W + RC_OFF_W0, 48, mb * 16u, 16u, lane),
                             xin[0], xin[1], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            acc[mb] = mfma16(a_rows(W + RC_OFF_W0, 48, mb * 16u, 32u, lane), xin[2], acc[mb]);
        }
        for(uint32_t mb = 0; mb < 4; mb++) h[0][mb] = relu_bf16(acc[mb]);
        for(uint32_t layer = 1; layer < 3; layer++)
        {
            const uint16_t* Wl = W + (layer == 1 ? RC_OFF_W1 : RC_OFF_W2);
            for(uint32_t mb = 0; mb < 4; mb++)
            {
                acc[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for(uint32_t ks = 0; ks < 4; ks += 2)
                    acc[mb] = mfma32(a_rows(Wl, 64, mb * 16u, ks * 16u, lane), a_rows(Wl, 64, mb * 16u, ks * 16u + 16u, lane),
                                     h[layer - 1][ks], h[layer - 1][ks + 1], acc[mb]);
            }
            for(uint32_t mb = 0; mb < 4; mb++) h[layer][mb] = relu_bf16(acc[mb]);
        }
        f32x4 y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for(uint32_t ks = 0; ks < 4; ks += 2)
            y = mfma32(a_rows(W + RC_OFF_W3, 64, 0, ks * 16u, lane), a_rows(W + RC_OFF_W3, 64, 0, ks * 16u + 16u, lane),
                       h[2][ks], h[2][ks + 1], y);
        // y: lane < 16 holds outputs 0..3 of sample s0 + lane (3 used)
        if(!TRAIN)
        {
            if(lane < 16u && sm < n)
            {
                // a row may carry an exactly computed term (direct light) in 13..15
                const float* sp = samples ? samples + (size_t)sm * RC_SAMPLE : nullptr;
                const float valid = sp ? sp[12] : 1.0f;
                const float b0 = sp ? sp[13] : 0.0f, b1 = sp ? sp[14] : 0.0f, b2 = sp ? sp[15] : 0.0f;
                reinterpret_cast<float4*>(pred)[sm] = make_float4(y[0] + b0, y[1] + b1, y[2] + b2, valid);
            }
            continue;
        }
        for(uint32_t ks = 0; ks < 3; ks++) store_T(acts.x0T, RC_IN, tile, ks, xin[ks], lane);
        for(uint32_t layer = 0; layer < 3; layer++)
            for(uint32_t mb = 0; mb < 4; mb++) store_T(acts.hT[layer], RC_W, tile, mb, h[layer][mb], lane);
        // dL/dy = (y - t) / B for valid samples
        f32x4 dy = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if(lane < 16u)
        {
            const float* sp = samples + (size_t)sm * RC_SAMPLE;
            if(sp[12] != 0.0f)
                for(int i = 0; i < 3; i++) dy[i] = (y[i] - sp[9 + i]) * invB;
        }
        bf16x4 dyb;
        for(int i = 0; i < 4; i++) dyb[i] = (short)f2bf(dy[i]);
        store_T(acts.dyT, RC_OUT, tile, 0, dyb, lane);
        // dH3^T = W3^T dY^T (K = 16 outputs), then through the ReLUs
        bf16x4 dz[4];
        for(uint32_t mb = 0; mb < 4; mb++)
        {
            f32x4 d = mfma16(a_rows(WT + RC_OFF_W3, 16, mb * 16u, 0, lane), dyb, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            dz[mb] = relu_grad(d, h[2][mb]);
            store_T(acts.dzT[2], RC_W, tile, mb, dz[mb], lane);
        }
        for(int layer = 1; layer >= 0; layer--)
        {
            // dH_layer^T = W_{layer+1}^T dZ_{layer+1}^T
            const uint16_t* WTl = WT + (layer == 1 ? RC_OFF_W2 : RC_OFF_W1);
            bf16x4 nz[4];
            for(uint32_t mb = 0; mb < 4; mb++)
            {
                f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for(uint32_t ks = 0; ks < 4; ks += 2)
                    d = mfma32(a_rows(WTl, 64, mb * 16u, ks * 16u, lane), a_rows(WTl, 64, mb * 16u, ks * 16u + 16u, lane),
                               dz[ks], dz[ks + 1], d);
                nz[mb] = relu_grad(d, h[layer][mb]);
                store_T(acts.dzT[layer], RC_W, tile, mb, nz[mb], lane);
            }
            for(uint32_t mb = 0; mb < 4; mb++) dz[mb] = nz[mb];
        }
        // dX0^T = W0^T dZ1^T, encoded features only (rows 0..31)
        for(uint32_t mb = 0; mb < 2; mb++)
        {
            f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            for(uint32_t ks = 0; ks < 4; ks += 2)
                d = mfma32(a_rows(WT + RC_OFF_W0, 64, mb * 16u, ks * 16u, lane),
                           a_rows(WT + RC_OFF_W0, 64, mb * 16u, ks * 16u + 16u, lane), dz[ks], dz[ks + 1], d);
            reinterpret_cast<float4*>(acts.denc + (size_t)sm * RC_ENC + mb * 16u + 4u * (lane >> 4))[0] =
                make_float4(d[0], d[1], d[2], d[3]);
        }
    }
}

// Weight gradients dW = dZ^T X summed over all tiles (K = samples, one
// 16-sample tile per v_mfma_f32_16x16x16_bf16: pairing two tiles into one
// K = 32 step doubles the scattered 8-byte operand loads per step and measured
// 153 against 69 us per 2^17 samples at 128 blocks); blocks take contiguous tile ranges; wave w owns output rows 16w..16w+15 of W0..W2 and output
// column block w of W3; partial sums go out by FP32 atomics.
__global__ __launch_bounds__(256) void k_rc_wgrad(DNrcNet net, uint32_t tiles, DNrcActs acts)
{
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t per = (tiles + gridDim.x - 1u) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per, t1 = min(tiles, t0 + per);
    f32x4 a0[3], a1[4], a2[4], a3;
    for(int i = 0; i < 3; i++) a0[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for(int i = 0; i < 4; i++) { a1[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f}; a2[i] = a1[i]; }
    a3 = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t row = w * 16u + (lane & 15u);           // A rows: output features
    const uint32_t kq = 4u * (lane >> 4);                  // samples 4*(lane>>4)..+3 of the tile
    for(uint32_t t = t0; t < t1; t++)
    {
        const size_t b64 = (size_t)t * RC_W * 16u, b48 = (size_t)t * RC_IN * 16u, b16 = (size_t)t * RC_OUT * 16u;
        bf16x4 A0 = ld_bf16x4(acts.dzT[0] + b64 + row * 16u + kq);
        bf16x4 A1 = ld_bf16x4(acts.dzT[1] + b64 + row * 16u + kq);
        bf16x4 A2 = ld_bf16x4(acts.dzT[2] + b64 + row * 16u + kq);
        for(uint32_t nb = 0; nb < 3; nb++)
            a0[nb] = mfma16(A0, ld_bf16x4(acts.x0T + b48 + (nb * 16u + (lane & 15u)) * 16u + kq), a0[nb]);
        for(uint32_t nb = 0; nb < 4; nb++)
        {
            const uint32_t col = (nb * 16u + (lane & 15u)) * 16u + kq;
            a1[nb] = mfma16(A1, ld_bf16x4(acts.hT[0] + b64 + col), a1[nb]);
            a2[nb] = mfma16(A2, ld_bf16x4(acts.hT[1] + b64 + col), a2[nb]);
        }
        bf16x4 A3 = ld_bf16x4(acts.dyT + b16 + (lane & 15u) * 16u + kq);
        a3 = mfma16(A3, ld_bf16x4(acts.hT[2] + b64 + (w * 16u + (lane & 15u)) * 16u + kq), a3);
    }
    if(t0 >= t1) return;
    // C layout: column (input feature) nb*16 + (lane & 15), rows 16w + 4*(lane>>4) + i
    const uint32_t r0 = w * 16u + 4u * (lane >> 4);
    for(uint32_t i = 0; i < 4; i++)
    {
        for(uint32_t nb = 0; nb < 3; nb++)
            atomicAdd(net.wGrad + RC_OFF_W0 + (r0 + i) * 48u + nb * 16u + (lane & 15u), a0[nb][i]);
        for(uint32_t nb = 0; nb < 4; nb++)
        {
            atomicAdd(net.wGrad + RC_OFF_W1 + (r0 + i) * 64u + nb * 16u + (lane & 15u), a1[nb][i]);
            atomicAdd(net.wGrad + RC_OFF_W2 + (r0 + i) * 64u + nb * 16u + (lane & 15u), a2[nb][i]);
        }
        // W3: rows = outputs 4*(lane>>4) + i, column block w
        atomicAdd(net.wGrad + RC_OFF_W3 + (4u * (lane >> 4) + i) * 64u + w * 16u + (lane & 15u), a3[i]);
    }
}

// Adam on FP32 master copies; refreshes the bf16 copies and clears the
// gradients. Weights: one thread per parameter; table: one per entry (2).
__device__ __forceinline__ float adam_one(float& w, float g, float& m, float& v, const DNrcAdam& o)
{
    m = o.beta1 * m + (1.0f - o.beta1) * g;
    v = o.beta2 * v + (1.0f - o.beta2) * g * g;
    w = w - o.lr * (m / o.bias1) / (sqrtf(v / o.bias2) + o.eps);
    return w;
}

__device__ __forceinline__ float ema_one(float& e, float w, const DNrcAdam& o)
{
    e = o.emaDecay * e + (1.0f - o.emaDecay) * w;
    return e * o.emaScale;
}

__global__ __launch_bounds__(256) void k_rc_adam(DNrcNet net, DNrcAdam o, size_t entries)
{
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if(i < RC_NPARAM)
    {
        float w = net.w32[i], m = net.wM[i], v = net.wV[i], e = net.wE[i];
        adam_one(w, net.wGrad[i], m, v, o);
        net.w32[i] = w; net.wM[i] = m; net.wV[i] = v;
        net.wGrad[i] = 0.0f;
        net.w16[i] = f2bf(w);
        net.wT16[rc_transposed_index((uint32_t)i)] = f2bf(w);
        net.wE16[i] = f2bf(ema_one(e, w, o));
        net.wE[i] = e;
    }
    if(i < entries)
    {
        float2 w = reinterpret_cast<float2*>(net.tab32)[i];
        float2 g = reinterpret_cast<float2*>(net.tabGrad)[i];
        float2 m = reinterpret_cast<float2*>(net.tabM)[i];
        float2 v = reinterpret_cast<float2*>(net.tabV)[i];
        float2 e = reinterpret_cast<float2*>(net.tabE)[i];
        adam_one(w.x, g.x, m.x, v.x, o);
        adam_one(w.y, g.y, m.y, v.y, o);
        reinterpret_cast<float2*>(net.tab32)[i] = w;
        reinterpret_cast<float2*>(net.tabM)[i] = m;
        reinterpret_cast<float2*>(net.tabV)[i] = v;
        reinterpret_cast<float2*>(net.tabGrad)[i] = make_float2(0.0f, 0.0f);
        net.tab16[i] = (uint32_t)f2bf(w.x) | ((uint32_t)f2bf(w.y) << 16);
        const float ex = ema_one(e.x, w.x, o), ey = ema_one(e.y, w.y, o);
        reinterpret_cast<float2*>(net.tabE)[i] = e;
        net.tabE16[i] = (uint32_t)f2bf(ex) | ((uint32_t)f2bf(ey) << 16);
    }
}

// ---------------------------------------------------------------------------
// training samples: one per lane, a stored voxel of the request list drawn by
// a per-sample hash, lit by the lighting pass's own device functions with
// per-sample random numbers (PCG hash, Jarzynski & Olano, "Hash Functions for
// GPU Rendering", JCGT 2020) instead of the frame-uniform ones.

__device__ __forceinline__ uint32_t pcg(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

struct SampleRandom {
    uint32_t seed;
    __device__ float u(uint32_t stream, uint32_t k) const
    {
        return (float)(pcg(seed ^ pcg(stream * 0x9E3779B9u + k)) >> 8) * (1.0f / 16777216.0f);
    }
    // uniform in the unit ball by rejection from the cube (as rand_unit_sphere)
    __device__ f3 ball(uint32_t stream) const
    {
        f3 p = splat(0.0f);
        for(uint32_t t = 0; t < 16u; t++)
        {
            p = mk(u(stream, 3 * t) * 2.0f - 1.0f, u(stream, 3 * t + 1) * 2.0f - 1.0f, u(stream, 3 * t + 2) * 2.0f - 1.0f);
            if(dot3(p, p) < 1.0f) return p;
        }
        return p * 0.5f;
    }
    __device__ float glossy_pick(uint32_t sample, uint32_t i) const { return u(100u + sample * 16u + i, 0) * 2.0f - 1.0f; }
    __device__ f3 glossy_dir(uint32_t sample, uint32_t i) const { return ball(200u + sample * 16u + i); }
    __device__ f3 diffuse_dir(uint32_t sample, uint32_t i) const { return ball(300u + sample * 16u + i); }
};

// Locates request slot `slot` (r * 32 + lane): false for an empty lane.
__device__ __forceinline__ bool slot_voxel(const DNlightParams& P, uint32_t slot, const DNchunkRec*& rec,
                                           uint32_t& mapIndex, uint32_t& voxNum, int32_t& sel)
{
    const uint32_t req = P.requests[slot >> 5];
    mapIndex = req >> 4;
    voxNum = (slot & 31u) + (req & 15u) * 32u;
    rec = P.scene.chunks + mapIndex;
    if(voxNum >= rec->numVoxels) return false;
    sel = select_in(rec->mask, voxNum);
    return sel >= 0;
}

__device__ __forceinline__ void write_sample(float* out, f3 pos, const DVoxel& me, f3 target, float valid,
                                             f3 exact = f3{0.0f, 0.0f, 0.0f})
{
    float4* o = reinterpret_cast<float4*>(out);
    o[0] = make_float4(pos.x, pos.y, pos.z, me.normal.x);
    o[1] = make_float4(me.normal.y, me.normal.z, me.albedo.x, me.albedo.y);
    o[2] = make_float4(me.albedo.z, target.x, target.y, target.z);
    o[3] = make_float4(valid, exact.x, exact.y, exact.z);
}

// What the cache learns and what is computed exactly. With fixed direct rays
// (numDirectRays > 0) the ambient and direct terms of a voxel's estimate
// (voxelLighting.comp:255,260) are deterministic: they are evaluated exactly
// and the network learns only the path-traced indirect term. With the
// jittered shadow ray the whole estimate is noisy and the network learns all
// of it.
__device__ __forceinline__ bool direct_is_exact(const DNlightParams& P) { return P.numDirectRays != 0u; }

struct VoxelSite {
    const DNchunkRec* rec;
    uint32_t voxNum;
    int32_t sel;
    DVoxel me;
    f3 p;      // ray origin, voxelLighting.comp:231-232
    f3 pos;    // voxel centre in voxel units
};

__device__ __forceinline__ VoxelSite voxel_site(const DNsceneView& s, const DNchunkRec* rec, uint32_t voxNum, int32_t sel)
{
    VoxelSite v;
    v.rec = rec; v.voxNum = voxNum; v.sel = sel;
    v.me = decompress(s.voxels[rec->voxelIndex + voxNum]);
    const int32_t cx = sel % 8, cy = (sel / 8) % 8, cz = sel / 64;
    v.p = (mk((float)cx, (float)cy, (float)cz) * 0.125f + mk((float)rec->pos[0], (float)rec->pos[1], (float)rec->pos[2])) +
          splat(0.0625f);
    v.p = v.p + splat(0.0625f - DN_EPS) * v.me.normal;
    v.pos = mk((float)(rec->pos[0] * 8 + cx), (float)(rec->pos[1] * 8 + cy), (float)(rec->pos[2] * 8 + cz)) + splat(0.5f);
    return v;
}

__device__ __forceinline__ void site_ctx(const DNlightParams& P, const DNsceneView& s, RayCtx& c, LCtx& L,
                                         uint32_t* maskRows)
{
    ctx_init(c);
    c.maskRow = lane_mask_row(maskRows);
    L.P = &P; L.s = &s; L.c = &c; L.dry = true; L.firstSample = false;
    L.sunDir = ld3(P.sunDir);
    L.sunStrength = ld3(s.sunStrength);
    L.direct = nullptr;
}

// Lane G draws sample g = G mod n of batch k = G / n, with that batch's seed
// seed0 + k * seedStride (the batches of several training steps in one launch).
__global__ __launch_bounds__(256) void k_rc_samples(DNlightParams P, uint32_t n, uint32_t total, uint32_t seed0,
                                                   uint32_t seedStride, float* samples)
{
    __shared__ uint32_t maskRows[DN_MASK_LDS_WORDS(256)];
    const uint32_t G = blockIdx.x * 256u + threadIdx.x;
    if(G >= total) return;
    const uint32_t k = G / n, g = G - k * n;
    const uint32_t seed = seed0 + k * seedStride;
    const uint32_t slots = *P.numRequests * 32u;
    const uint32_t base = pcg(seed ^ pcg(g));
    const DNchunkRec* rec = nullptr;
    uint32_t mapIndex = 0, voxNum = 0;
    int32_t sel = -1;
    bool found = false;
    // first try: slot g & 31 of a request drawn per run of 32 samples (the
    // runs keep a wave's samples in two chunks); then random slots. Over the
    // 32 positions of a run every stored voxel is as likely as any other.
    for(uint32_t t = 0; t < 8u && slots && !found; t++)
    {
        const uint32_t slot = t == 0 ? (pcg(seed ^ pcg((g >> 5) ^ 0x5BD1E995u)) % (slots >> 5)) * 32u + (g & 31u)
                                     : pcg(base + t * 0x68E31DA4u) % slots;
        found = slot_voxel(P, slot, rec, mapIndex, voxNum, sel);
    }
    float* out = samples + (size_t)G * RC_SAMPLE;
    if(!found)
    {
        write_sample(out, splat(0.0f), DVoxel{}, splat(0.0f), 0.0f);
        return;
    }
    const DNsceneView& s = P.scene;
    const VoxelSite v = voxel_site(s, rec, voxNum, sel);
    const DNmat& mat = s.mats[v.me.material];
    RayCtx c;
    LCtx L;
    site_ctx(P, s, c, L, maskRows);
    f3 diff = splat(0.0f);
    if(mat.specular < 1.0f)
    {
        const SampleRandom rng{base};
        diffuse_ray(L, v.me.normal + DN_EPS, v.p, v.me, 0, diff, rng);
        if(!direct_is_exact(P))
        {
            diff = diff + ld3(P.ambient);
            f3 dir = vnorm(L.sunDir * P.shadowSoftness + rng.ball(400u)) + DN_EPS;
            DVoxel hv = {};
            f3 ca; float cm;
            f3 pp = v.p;
            if(!light_trace(s, c, dir, vinv(dir), pp, hv, ca, cm)) diff = diff + (L.sunStrength * cm + ca);
        }
    }
    write_sample(out, v.pos, v.me, diff, 1.0f);
}

// Query rows for every slot of the request list (valid 0 for empty lanes).
__global__ __launch_bounds__(256) void k_rc_gather(DNlightParams P, uint32_t slotsCap, float* samples)
{
    __shared__ uint32_t maskRows[DN_MASK_LDS_WORDS(256)];
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if(g >= slotsCap) return;
    float* out = samples + (size_t)g * RC_SAMPLE;
    const DNchunkRec* rec = nullptr;
    uint32_t mapIndex = 0, voxNum = 0;
    int32_t sel = -1;
    if(g >= *P.numRequests * 32u || !slot_voxel(P, g, rec, mapIndex, voxNum, sel))
    {
        write_sample(out, splat(0.0f), DVoxel{}, splat(0.0f), 0.0f);
        return;
    }
    const DNsceneView& s = P.scene;
    const VoxelSite v = voxel_site(s, rec, voxNum, sel);
    f3 exact = splat(0.0f);
    if(direct_is_exact(P) && s.mats[v.me.material].specular < 1.0f)
    {
        // ambient + the fixed direct rays, as the lighting pass adds them
        RayCtx c;
        LCtx L;
        site_ctx(P, s, c, L, maskRows);
        L.direct = (P.directIn && g < P.directSlots) ? P.directIn + g : nullptr;
        exact = ld3(P.ambient);
        shadow_ray(L, v.p, 0, exact);
    }
    write_sample(out, v.pos, v.me, splat(0.0f), 1.0f, exact);
}

// Reference estimate of every slot's irradiance: `per` independent samples per
// thread (per-sample random numbers), `chunks` threads per slot, summed in
// FP64 into acc[slot] (rgb, sample count). With fixed direct rays each sample
// is the exact ambient + direct term plus one path-traced indirect term.
__global__ __launch_bounds__(256) void k_rc_reference(DNlightParams P, uint32_t slotsCap, uint32_t chunks, uint32_t per,
                                                      uint32_t seed, double* acc)
{
    __shared__ uint32_t maskRows[DN_MASK_LDS_WORDS(256)];
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    const uint32_t slot = g / chunks, chunk = g % chunks;
    if(slot >= slotsCap || slot >= *P.numRequests * 32u) return;
    const DNchunkRec* rec = nullptr;
    uint32_t mapIndex = 0, voxNum = 0;
    int32_t sel = -1;
    if(!slot_voxel(P, slot, rec, mapIndex, voxNum, sel)) return;
    const DNsceneView& s = P.scene;
    const VoxelSite v = voxel_site(s, rec, voxNum, sel);
    if(s.mats[v.me.material].specular >= 1.0f) return;
    RayCtx c;
    LCtx L;
    site_ctx(P, s, c, L, maskRows);
    const bool exactDirect = direct_is_exact(P);
    f3 exact = splat(0.0f);
    if(exactDirect)
    {
        exact = ld3(P.ambient);
        shadow_ray(L, v.p, 0, exact);
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for(uint32_t k = 0; k < per; k++)
    {
        const SampleRandom rng{pcg(seed ^ pcg(slot * 0x9E3779B9u ^ pcg(chunk * per + k)))};
        f3 smp = splat(0.0f);
        diffuse_ray(L, v.me.normal + DN_EPS, v.p, v.me, 0, smp, rng);
        if(exactDirect)
            smp = smp + exact;
        else
        {
            smp = smp + ld3(P.ambient);
            f3 dir = vnorm(L.sunDir * P.shadowSoftness + rng.ball(400u)) + DN_EPS;
            DVoxel hv = {};
            f3 ca; float cm;
            f3 pp = v.p;
            if(!light_trace(s, c, dir, vinv(dir), pp, hv, ca, cm)) smp = smp + (L.sunStrength * cm + ca);
        }
        sx += (double)smp.x; sy += (double)smp.y; sz += (double)smp.z;
    }
    double* a = acc + (size_t)slot * 4u;
    atomicAdd(a + 0, sx);
    atomicAdd(a + 1, sy);
    atomicAdd(a + 2, sz);
    atomicAdd(a + 3, (double)per);
}

__global__ __launch_bounds__(256) void k_accumulate_f64(const float* src, double* dst, size_t n)
{
    for(size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u)
        dst[i] += (double)src[i];
}

inline uint32_t blocks_for(size_t threads, uint32_t cap = 0x7FFFFFFFu)
{
    size_t b = (threads + 255u) / 256u;
    if(b < 1) b = 1;
    return (uint32_t)(b > cap ? cap : b);
}

}  // namespace

int dn_rc_encode(const DNrcNet& net, const float* samples, uint32_t n, uint16_t* x0, hipStream_t stream)
{
    if(n == 0) return 0;
    hipLaunchKernelGGL(k_rc_encode, dim3(blocks_for((size_t)n * 16u)), dim3(256), 0, stream, net, samples, n, x0);
    return (int)hipGetLastError();
}

int dn_rc_forward(const DNrcNet& net, const uint16_t* x0, const float* samples, uint32_t n, float* pred,
                  hipStream_t stream)
{
    if(n == 0 || (n & 15u)) return n ? (int)hipErrorInvalidValue : 0;
    DNrcActs none = {};
    hipLaunchKernelGGL(k_rc_mlp<false>, dim3(blocks_for((size_t)n / 16u * 64u, 1024)), dim3(256), 0, stream, net, x0,
                       samples, n, 0.0f, pred, none);
    return (int)hipGetLastError();
}

int dn_rc_train(const DNrcNet& net, const float* samples, uint32_t n, const DNrcActs& a, hipStream_t stream,
                const DNrcSide* side)
{
    if(n == 0 || (n & 15u)) return n ? (int)hipErrorInvalidValue : 0;
    const uint32_t tiles = n / 16u;
    hipLaunchKernelGGL(k_rc_encode, dim3(blocks_for((size_t)n * 16u)), dim3(256), 0, stream, net, samples, n, a.x0);
    // at most 512 blocks (4 tiles per wave at 2^17 samples): 83 us per 2^17 samples against 92 at
    // 1024 blocks, 98 at 256 and 82 at 768
    hipLaunchKernelGGL(k_rc_mlp<true>, dim3(blocks_for((size_t)tiles * 64u, 512)), dim3(256), 0, stream, net,
                       (const uint16_t*)a.x0, samples, n, 1.0f / (float)n, (float*)nullptr, a);
    // The weight gradients (MFMA and LDS bound) and the table gradients (bound
    // by their atomics) read the same activations and write disjoint buffers:
    // with a side stream they run side by side.
    // one block per CU (47 us per 2^17 samples; 128 blocks 71 us, 512 blocks 50 us)
    const uint32_t wb = tiles < 256u ? tiles : 256u;
    hipStream_t ws = stream;
    if(side)
    {
        if(hipEventRecord(side->fork, stream) != hipSuccess ||
           hipStreamWaitEvent(side->stream, side->fork, 0) != hipSuccess)
            return (int)hipErrorLaunchFailure;
        ws = side->stream;
    }
    hipLaunchKernelGGL(k_rc_wgrad, dim3(wb), dim3(256), 0, ws, net, tiles, a);
    // one wave per 16 samples and level
    hipLaunchKernelGGL(k_rc_table_grad, dim3(blocks_for((size_t)RC_LEVELS * ((n + 15u) / 16u) * 64u)), dim3(256), 0,
                       stream, net, samples, n, (const float*)a.denc);
    if(side && (hipEventRecord(side->join, side->stream) != hipSuccess ||
                hipStreamWaitEvent(stream, side->join, 0) != hipSuccess))
        return (int)hipErrorLaunchFailure;
    return (int)hipGetLastError();
}

int dn_rc_adam(const DNrcNet& net, const DNrcAdam& opt, hipStream_t stream)
{
    const size_t entries = (size_t)RC_LEVELS << net.log2T;
    hipLaunchKernelGGL(k_rc_adam, dim3(blocks_for(entries > RC_NPARAM ? entries : RC_NPARAM)), dim3(256), 0, stream,
                       net, opt, entries);
    return (int)hipGetLastError();
}

int dn_rc_samples(const DNlightParams& p, uint32_t maxRequests, uint32_t n, uint32_t batches, uint32_t seed,
                  uint32_t seedStride, float* samples, hipStream_t stream)
{
    (void)maxRequests;
    if(n == 0 || batches == 0) return 0;
    if((uint64_t)n * batches > 0xFFFFFFFFull) return (int)hipErrorInvalidValue;
    const uint32_t total = n * batches;
    hipLaunchKernelGGL(k_rc_samples, dim3(blocks_for(total)), dim3(256), 0, stream, p, n, total, seed, seedStride,
                       samples);
    return (int)hipGetLastError();
}

int dn_rc_gather(const DNlightParams& p, uint32_t slots, float* samples, hipStream_t stream)
{
    if(slots == 0) return 0;
    hipLaunchKernelGGL(k_rc_gather, dim3(blocks_for(slots)), dim3(256), 0, stream, p, slots, samples);
    return (int)hipGetLastError();
}

int dn_rc_reference(const DNlightParams& p, uint32_t slots, uint32_t samplesPerVoxel, uint32_t seed, double* acc,
                    hipStream_t stream)
{
    if(slots == 0 || samplesPerVoxel == 0) return 0;
    const uint32_t chunks = samplesPerVoxel >= 256u ? 16u : 1u;
    const uint32_t per = (samplesPerVoxel + chunks - 1u) / chunks;
    hipLaunchKernelGGL(k_rc_reference, dim3(blocks_for((size_t)slots * chunks)), dim3(256), 0, stream, p, slots, chunks,
                       per, seed, acc);
    return (int)hipGetLastError();
}

int dn_accumulate_f64(const float* src, double* dst, size_t n, hipStream_t stream)
{
    if(n == 0) return 0;
    hipLaunchKernelGGL(k_accumulate_f64, dim3(blocks_for(n, 4096)), dim3(256), 0, stream, src, dst, n);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free prediction (DNX_rc_predict_samples_noalloc). The encoded
// input rows that DNX_rc_predict_samples keeps in a query buffer grown with
// hipMalloc live here in `scratch`, a fixed array that every device holds in
// its own memory for the life of the process: DNX_RC_NOALLOC_MAX_SAMPLES rows
// of RC_IN bf16 values (24 MiB). k_rc_encode writes every column of the rows
// it is given and k_rc_mlp reads only those, so nothing needs clearing between
// calls. In host code `scratch` names a host-side stand-in, not device memory:
// kernels get the address that hipGetSymbolAddress gives for each device.
static_assert(DNX_RC_NOALLOC_MAX_SAMPLES % 16u == 0u, "whole 16-sample tiles");
static __device__ __align__(256) uint16_t scratch[(size_t)DNX_RC_NOALLOC_MAX_SAMPLES * RC_IN];

namespace {
std::mutex g_scratchLock;
std::vector<uint16_t*> g_scratchAddr;     // per device: scratch's device address
}  // namespace

bool dn_noalloc_init()
{
    int count = 0, current = 0;
    if(hipGetDeviceCount(&count) != hipSuccess || count < 1 || hipGetDevice(&current) != hipSuccess) return false;
    std::lock_guard<std::mutex> lock(g_scratchLock);
    if(g_scratchAddr.size() < (size_t)count) g_scratchAddr.resize((size_t)count, nullptr);
    bool ok = true;
    for(int d = 0; d < count; d++)
    {
        if(g_scratchAddr[d]) continue;                       // once per device
        void* p = nullptr;
        if(hipSetDevice(d) != hipSuccess || hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        {
            ok = false;
            continue;
        }
        g_scratchAddr[d] = static_cast<uint16_t*>(p);
    }
    if(hipSetDevice(current) != hipSuccess) ok = false;
    return ok;
}

int dn_rc_predict_noalloc(const DNrcNet& net, int device, const float* samples, uint32_t n, float* pred,
                          hipStream_t stream)
{
    if(n == 0 || (n & 15u) || n > DNX_RC_NOALLOC_MAX_SAMPLES) return (int)hipErrorInvalidValue;
    uint16_t* x0 = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratchLock);
        if(device >= 0 && (size_t)device < g_scratchAddr.size()) x0 = g_scratchAddr[device];
    }
    if(!x0) return (int)hipErrorNotInitialized;
    const int e = dn_rc_encode(net, samples, n, x0, stream);
    return e ? e : dn_rc_forward(net, x0, samples, n, pred, stream);
}

// Internal layouts and kernel interfaces of the radiance cache (rcache.cpp,
// rcache.hip), apart from the frame pipeline's in dn_internal.h. Nothing here
// is part of the public C-ABI.
#pragma once
#include "dn_internal.h"

// Radiance cache network: multiresolution hash encoding (RC_LEVELS levels of
// RC_FEAT features) plus an MLP of four weight matrices (3 hidden layers of
// RC_W with ReLU, linear output of 3), no biases. Weights are stored row-major
// [out][in] and zero-padded: W0 64x48 (38 inputs used), W1 64x64, W2 64x64,
// W3 16x64 (3 outputs used). A sample row is 16 floats: position (voxel
// units) 3, normal 3, albedo 3, target 3, valid (0/1), 3 unused.
enum {
    RC_LEVELS = 16, RC_FEAT = 2, RC_ENC = 32, RC_IN = 48, RC_W = 64, RC_OUT = 16, RC_SAMPLE = 16,
    RC_OFF_W0 = 0, RC_OFF_W1 = 3072, RC_OFF_W2 = 7168, RC_OFF_W3 = 11264, RC_NPARAM = 12288
};

// Index in the transposed weight layout (each of W0 64x48, W1 and W2 64x64,
// W3 16x64 transposed in place, same offsets) of weight i of the row-major one.
__host__ __device__ inline uint32_t rc_transposed_index(uint32_t i)
{
    uint32_t off, rows, cols;
    if(i < RC_OFF_W1) { off = RC_OFF_W0; rows = 64; cols = 48; }
    else if(i < RC_OFF_W2) { off = RC_OFF_W1; rows = 64; cols = 64; }
    else if(i < RC_OFF_W3) { off = RC_OFF_W2; rows = 64; cols = 64; }
    else { off = RC_OFF_W3; rows = 16; cols = 64; }
    const uint32_t j = i - off;
    return off + (j % cols) * rows + j / cols;
}

struct DNrcNet {
    uint32_t log2T;
    uint32_t denseMask;            // bit l: level l indexes densely ((N_l + 2)^3 <= T)
    uint32_t res[RC_LEVELS];       // N_l
    float extent;                  // positions / extent in [0, 1)
    uint32_t* tab16;               // [L][T] two bf16 features per entry
    float* tab32;                  // [L][T][2] master copy
    float* tabGrad;
    float* tabM;
    float* tabV;
    uint16_t* w16;                 // [RC_NPARAM] bf16
    uint16_t* wT16;                // w16 in the transposed layout (rc_transposed_index), for backward
    float* w32;                    // master copy
    float* wGrad;
    float* wM;
    float* wV;
    // Exponential moving average of the parameters over the Adam steps, and
    // its debiased bf16 copies, which predictions use.
    float* tabE;
    uint32_t* tabE16;
    float* wE;
    uint16_t* wE16;
};
// Per-batch activation buffers of training, `tiles` = samples / 16. The *T
// buffers are tile-transposed bf16: [tile][feature][16 samples].
struct DNrcActs {
    uint16_t* x0;                  // [samples][48] bf16, sample-major
    uint16_t* x0T;                 // [tile][48][16]
    uint16_t* hT[3];               // [tile][64][16]
    uint16_t* dzT[3];              // [tile][64][16]
    uint16_t* dyT;                 // [tile][16][16]
    float* denc;                   // [samples][32]
};
struct DNrcAdam {
    float lr, beta1, beta2, eps;
    float bias1, bias2;            // 1 - beta^t
    float emaDecay;                // d: E = d * E + (1 - d) * W after each step
    float emaScale;                // 1 / (1 - d^t): debiasing of E (which starts at 0)
};

int dn_rc_encode(const DNrcNet& net, const float* samples, uint32_t n, uint16_t* x0, hipStream_t stream);
// pred: float4 per row (3 outputs, the row's valid flag; 1 without samples).
int dn_rc_forward(const DNrcNet& net, const uint16_t* x0, const float* samples, uint32_t n, float* pred,
                  hipStream_t stream);
// A second stream for training: the weight gradients run on it beside the
// table gradients (`fork` and `join` order it after the MLP and before Adam).
struct DNrcSide {
    hipStream_t stream;
    hipEvent_t fork, join;
};
// side: null runs every kernel on `stream`.
int dn_rc_train(const DNrcNet& net, const float* samples, uint32_t n, const DNrcActs& a, hipStream_t stream,
                const DNrcSide* side = nullptr);
// Allocation-free prediction: dn_rc_encode + dn_rc_forward with the encoded rows
// in the fixed per-device scratch array of rcache.hip (at most
// DNX_RC_NOALLOC_MAX_SAMPLES rows) on device `device`. dn_noalloc_init (called
// by DN_init) looks up every device's address of that array first.
bool dn_noalloc_init();
int dn_rc_predict_noalloc(const DNrcNet& net, int device, const float* samples, uint32_t n, float* pred,
                          hipStream_t stream);
int dn_rc_adam(const DNrcNet& net, const DNrcAdam& opt, hipStream_t stream);
// `batches` batches of n training samples, batch k drawn with seed + k * seedStride
int dn_rc_samples(const DNlightParams& p, uint32_t maxRequests, uint32_t n, uint32_t batches, uint32_t seed,
                  uint32_t seedStride, float* samples, hipStream_t stream);
int dn_rc_gather(const DNlightParams& p, uint32_t slots, float* samples, hipStream_t stream);
int dn_accumulate_f64(const float* src, double* dst, size_t n, hipStream_t stream);
// acc[slot] (4 doubles: rgb sum, count) += samplesPerVoxel (rounded up to the
// work split) independent samples of every stored voxel of the request list.
int dn_rc_reference(const DNlightParams& p, uint32_t slots, uint32_t samplesPerVoxel, uint32_t seed, double* acc,
                    hipStream_t stream);

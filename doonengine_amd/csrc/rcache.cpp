// Host side of the learned radiance cache (no counterpart in the reference):
// parameter storage, initialisation, the per-frame train / predict sequence
// and the C-ABI entry points declared in doon/doon.h.
#include "doon/doon.h"
#include "rc_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

void dn_msg(DNmessageType type, DNmessageSeverity sev, const char* fmt, ...);

struct DNXradianceCache {
    DNvolume* vol = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;               // the device its buffers live on (current at creation)
    DNXrcConfig cfg;
    DNrcNet net;
    DNrcActs acts;
    uint32_t cap = 0;             // rows of the training buffers
    float* dSamples = nullptr;    // sampleCap rows: the drawn training samples of up to 16 steps
    size_t sampleCap = 0;
    const float* lastSamplesPtr = nullptr;   // the last step's rows in dSamples
    uint32_t lastSamples = 0;
    float* dQuery = nullptr;      // query rows for predictions
    uint16_t* dQueryX0 = nullptr;
    size_t queryCap = 0;
    uint32_t adamStep = 0;
    double emaKeep = 1.0;         // product of the average's decays so far (its debiasing)
    uint32_t frame = 0;
    DNrcSide side = {nullptr, nullptr, nullptr};   // training's second stream (side.stream null: none)
};

namespace {

bool ok(hipError_t e, const char* what)
{
    if(e == hipSuccess) return true;
    dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR, "radiance cache: %s: %s", what, hipGetErrorString(e));
    return false;
}

template <typename T> bool dev(T** p, size_t count, const char* what)
{
    *p = nullptr;
    return ok(hipMalloc((void**)p, sizeof(T) * (count ? count : 1)), what);
}
template <typename T> void dfree(T*& p)
{
    if(p) hipFree(p);
    p = nullptr;
}

// splitmix64: the seeded generator of the initial parameters
struct Mix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
};

size_t table_floats(const DNXradianceCache* rc) { return (size_t)RC_LEVELS * 2u << rc->net.log2T; }

uint16_t host_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Uploads fp32 master parameters with their bf16 copies; clears gradients and
// Adam moments.
bool upload_params(DNXradianceCache* rc, const float* w, const float* tab)
{
    const size_t tf = table_floats(rc);
    std::vector<uint16_t> w16(RC_NPARAM), wT16(RC_NPARAM);
    for(int i = 0; i < RC_NPARAM; i++) w16[i] = host_bf16(w[i]);
    for(uint32_t i = 0; i < RC_NPARAM; i++) wT16[rc_transposed_index(i)] = w16[i];
    std::vector<uint32_t> t16(tf / 2);
    for(size_t i = 0; i < tf / 2; i++) t16[i] = host_bf16(tab[2 * i]) | ((uint32_t)host_bf16(tab[2 * i + 1]) << 16);
    DNrcNet& n = rc->net;
    bool good = ok(hipMemcpyAsync(n.w32, w, sizeof(float) * RC_NPARAM, hipMemcpyHostToDevice, rc->stream), "upload") &&
                ok(hipMemcpyAsync(n.w16, w16.data(), 2 * RC_NPARAM, hipMemcpyHostToDevice, rc->stream), "upload") &&
                ok(hipMemcpyAsync(n.wT16, wT16.data(), 2 * RC_NPARAM, hipMemcpyHostToDevice, rc->stream), "upload") &&
                ok(hipMemcpyAsync(n.tab32, tab, sizeof(float) * tf, hipMemcpyHostToDevice, rc->stream), "upload") &&
                ok(hipMemcpyAsync(n.tab16, t16.data(), 4 * (tf / 2), hipMemcpyHostToDevice, rc->stream), "upload");
    good = good && ok(hipMemcpyAsync(n.wE16, w16.data(), 2 * RC_NPARAM, hipMemcpyHostToDevice, rc->stream), "upload") &&
           ok(hipMemcpyAsync(n.tabE16, t16.data(), 4 * (tf / 2), hipMemcpyHostToDevice, rc->stream), "upload");
    for(float* p : {n.wGrad, n.wM, n.wV, n.wE})
        good = good && ok(hipMemsetAsync(p, 0, sizeof(float) * RC_NPARAM, rc->stream), "clear");
    for(float* p : {n.tabGrad, n.tabM, n.tabV, n.tabE})
        good = good && ok(hipMemsetAsync(p, 0, sizeof(float) * tf, rc->stream), "clear");
    good = good && ok(hipStreamSynchronize(rc->stream), "upload");
    rc->adamStep = 0;
    rc->emaKeep = 1.0;
    return good;
}

bool ensure_train(DNXradianceCache* rc, uint32_t n)
{
    if(n <= rc->cap) return true;
    DNrcActs& a = rc->acts;
    hipStreamSynchronize(rc->stream);
    dfree(a.x0); dfree(a.x0T); dfree(a.dyT); dfree(a.denc);
    for(int i = 0; i < 3; i++) { dfree(a.hT[i]); dfree(a.dzT[i]); }
    bool g = dev(&a.x0, (size_t)n * RC_IN, "activations") && dev(&a.x0T, (size_t)n * RC_IN, "activations") &&
             dev(&a.dyT, (size_t)n * RC_OUT, "activations") && dev(&a.denc, (size_t)n * RC_ENC, "activations");
    for(int i = 0; g && i < 3; i++)
        g = dev(&a.hT[i], (size_t)n * RC_W, "activations") && dev(&a.dzT[i], (size_t)n * RC_W, "activations");
    rc->cap = g ? n : 0;
    return g;
}

bool ensure_samples(DNXradianceCache* rc, size_t rows)
{
    if(rows <= rc->sampleCap) return true;
    hipStreamSynchronize(rc->stream);
    dfree(rc->dSamples);
    rc->lastSamplesPtr = nullptr;
    rc->lastSamples = 0;
    const bool g = dev(&rc->dSamples, rows * RC_SAMPLE, "samples");
    rc->sampleCap = g ? rows : 0;
    return g;
}

bool ensure_query(DNXradianceCache* rc, size_t n)
{
    if(n <= rc->queryCap) return true;
    hipStreamSynchronize(rc->stream);
    dfree(rc->dQuery);
    dfree(rc->dQueryX0);
    bool g = dev(&rc->dQuery, n * RC_SAMPLE, "queries") && dev(&rc->dQueryX0, n * RC_IN, "queries");
    rc->queryCap = g ? n : 0;
    return g;
}

bool adam_step(DNXradianceCache* rc)
{
    rc->adamStep++;
    DNrcAdam o;
    o.lr = rc->cfg.lr;
    if(rc->cfg.lrDecaySteps > 0.0f)
        o.lr = (float)((double)rc->cfg.lr / std::pow(1.0 + (double)(rc->adamStep - 1) / (double)rc->cfg.lrDecaySteps,
                                                     (double)rc->cfg.lrDecayPower));
    o.beta1 = rc->cfg.beta1;
    o.beta2 = rc->cfg.beta2;
    o.eps = rc->cfg.eps;
    o.bias1 = (float)(1.0 - std::pow((double)rc->cfg.beta1, (double)rc->adamStep));
    o.bias2 = (float)(1.0 - std::pow((double)rc->cfg.beta2, (double)rc->adamStep));
    // step t's decay: emaDecay, or while the window grows 1 - emaGrowth / t
    // (an average over about the last t / emaGrowth steps), at least 0. With
    // decays d_1..d_t the average's weights sum to 1 - d_1 ... d_t.
    double d = (double)rc->cfg.emaDecay;
    if(rc->cfg.emaGrowth > 0.0f) d = std::min(d, std::max(0.0, 1.0 - (double)rc->cfg.emaGrowth / rc->adamStep));
    rc->emaKeep *= d;
    o.emaDecay = (float)d;
    o.emaScale = (float)(1.0 / (1.0 - rc->emaKeep));
    return dn_rc_adam(rc->net, o, rc->stream) == 0;
}

// The parameters predictions read: the debiased moving average (whose bf16
// copies equal the parameters themselves until the first Adam step).
DNrcNet infer_net(const DNXradianceCache* rc)
{
    DNrcNet n = rc->net;
    n.tab16 = rc->net.tabE16;
    n.w16 = rc->net.wE16;
    return n;
}

}  // namespace

extern "C" void DNX_rc_default_config(DNXrcConfig* c)
{
    if(!c) return;
    c->log2Table = 19;
    c->nMin = 16;
    c->nMax = 256;
    c->extent = 0.0f;
    c->lr = 1e-2f;
    c->beta1 = 0.9f;
    c->beta2 = 0.99f;
    c->eps = 1e-10f;
    c->samplesPerFrame = 65536;
    c->seed = 1;
    c->emaDecay = 0.95f;
    c->lrDecaySteps = 0.0f;
    c->lrDecayPower = 0.5f;
    c->emaGrowth = 8.0f;
}

extern "C" void DNX_rc_destroy(DNXradianceCache* rc)
{
    if(!rc) return;
    if(rc->stream) hipStreamSynchronize(rc->stream);
    if(rc->side.stream)
    {
        hipStreamSynchronize(rc->side.stream);
        hipStreamDestroy(rc->side.stream);
        hipEventDestroy(rc->side.fork);
        hipEventDestroy(rc->side.join);
    }
    DNrcNet& n = rc->net;
    dfree(n.tab16); dfree(n.tab32); dfree(n.tabGrad); dfree(n.tabM); dfree(n.tabV); dfree(n.tabE); dfree(n.tabE16);
    dfree(n.w16); dfree(n.wT16); dfree(n.w32); dfree(n.wGrad); dfree(n.wM); dfree(n.wV); dfree(n.wE); dfree(n.wE16);
    DNrcActs& a = rc->acts;
    dfree(rc->dSamples); dfree(a.x0); dfree(a.x0T); dfree(a.dyT); dfree(a.denc);
    for(int i = 0; i < 3; i++) { dfree(a.hT[i]); dfree(a.dzT[i]); }
    dfree(rc->dQuery);
    dfree(rc->dQueryX0);
    delete rc;
}

extern "C" DNXradianceCache* DNX_rc_create(DNvolume* vol, const DNXrcConfig* config)
{
    if(!vol || !vol->gpu) return nullptr;
    DNXrcConfig c;
    DNX_rc_default_config(&c);
    if(config) c = *config;
    if(c.log2Table < 4 || c.log2Table > 24 || c.nMin < 1 || c.nMax < c.nMin || (c.samplesPerFrame & 15u) ||
       c.samplesPerFrame == 0 || !(c.lr > 0.0f) || !(c.emaDecay >= 0.0f && c.emaDecay < 1.0f) ||
       !(c.lrDecaySteps >= 0.0f) || !(c.lrDecayPower >= 0.0f) || !(c.emaGrowth >= 0.0f))
    {
        dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR,
               "radiance cache config: log2Table in [4, 24], 1 <= nMin <= nMax, samplesPerFrame a positive "
               "multiple of 16, lr > 0, emaDecay in [0, 1), lrDecaySteps, lrDecayPower, emaGrowth >= 0");
        return nullptr;
    }
    DNXradianceCache* rc = new DNXradianceCache();
    std::memset(&rc->net, 0, sizeof(rc->net));
    std::memset(&rc->acts, 0, sizeof(rc->acts));
    rc->vol = vol;
    rc->stream = (hipStream_t)DNX_get_stream(vol);
    // a second stream for training (default priority); without it training runs on one
    if(hipStreamCreateWithFlags(&rc->side.stream, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreateWithFlags(&rc->side.fork, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&rc->side.join, hipEventDisableTiming) != hipSuccess)
    {
        if(rc->side.stream) hipStreamDestroy(rc->side.stream);
        if(rc->side.fork) hipEventDestroy(rc->side.fork);
        if(rc->side.join) hipEventDestroy(rc->side.join);
        rc->side = DNrcSide{nullptr, nullptr, nullptr};
    }
    if(hipGetDevice(&rc->device) != hipSuccess) rc->device = 0;
    if(c.extent <= 0.0f)
    {
        uint32_t m = vol->mapSize.x > vol->mapSize.y ? vol->mapSize.x : vol->mapSize.y;
        m = m > vol->mapSize.z ? m : vol->mapSize.z;
        c.extent = 8.0f * (float)m;
    }
    rc->cfg = c;
    DNrcNet& n = rc->net;
    n.log2T = c.log2Table;
    n.extent = c.extent;
    const double b = std::exp((std::log((double)c.nMax) - std::log((double)c.nMin)) / (double)(RC_LEVELS - 1));
    for(int l = 0; l < RC_LEVELS; l++)
    {
        n.res[l] = (uint32_t)std::floor((double)c.nMin * std::pow(b, (double)l));
        const double s = (double)n.res[l] + 2.0;
        if(s * s * s <= (double)(1u << n.log2T)) n.denseMask |= 1u << l;
    }
    const size_t tf = table_floats(rc);
    bool g = dev(&n.tab16, tf / 2, "hash table") && dev(&n.tab32, tf, "hash table") && dev(&n.tabGrad, tf, "hash table") &&
             dev(&n.tabM, tf, "hash table") && dev(&n.tabV, tf, "hash table") && dev(&n.tabE, tf, "hash table") &&
             dev(&n.tabE16, tf / 2, "hash table") && dev(&n.w16, RC_NPARAM, "weights") && dev(&n.wT16, RC_NPARAM, "weights") &&
             dev(&n.w32, RC_NPARAM, "weights") && dev(&n.wGrad, RC_NPARAM, "weights") &&
             dev(&n.wM, RC_NPARAM, "weights") && dev(&n.wV, RC_NPARAM, "weights") && dev(&n.wE, RC_NPARAM, "weights") &&
             dev(&n.wE16, RC_NPARAM, "weights");
    if(g)
    {
        // Xavier-uniform weights on the used block of the hidden layers, a zero
        // output layer (the first predictions are 0, not a random scale);
        // table entries uniform in [-1e-4, 1e-4]
        Mix rng{0x5DEECE66Dull ^ ((uint64_t)c.seed << 32)};
        std::vector<float> w(RC_NPARAM, 0.0f);
        const struct { int off, ld, rows, cols; } mats[3] = {
            {RC_OFF_W0, 48, 64, 38}, {RC_OFF_W1, 64, 64, 64}, {RC_OFF_W2, 64, 64, 64}};
        for(const auto& m : mats)
        {
            const double a = std::sqrt(6.0 / (double)(m.rows + m.cols));
            for(int r = 0; r < m.rows; r++)
                for(int k = 0; k < m.cols; k++) w[m.off + r * m.ld + k] = (float)((rng.uniform() * 2.0 - 1.0) * a);
        }
        std::vector<float> tab(tf);
        for(size_t i = 0; i < tf; i++) tab[i] = (float)((rng.uniform() * 2.0 - 1.0) * 1e-4);
        g = upload_params(rc, w.data(), tab.data()) && ensure_train(rc, c.samplesPerFrame);
    }
    if(!g)
    {
        DNX_rc_destroy(rc);
        return nullptr;
    }
    return rc;
}

extern "C" bool DNX_rc_set_params(DNXradianceCache* rc, const float* weights, const float* table)
{
    if(!rc || !weights || !table) return false;
    return upload_params(rc, weights, table);
}

extern "C" bool DNX_rc_get_params(DNXradianceCache* rc, float* weights, float* table, float* weightGrad,
                                  float* tableGrad)
{
    if(!rc) return false;
    const size_t tf = table_floats(rc);
    bool g = true;
    if(weights) g = g && ok(hipMemcpyAsync(weights, rc->net.w32, 4 * RC_NPARAM, hipMemcpyDeviceToHost, rc->stream), "read");
    if(table) g = g && ok(hipMemcpyAsync(table, rc->net.tab32, 4 * tf, hipMemcpyDeviceToHost, rc->stream), "read");
    if(weightGrad)
        g = g && ok(hipMemcpyAsync(weightGrad, rc->net.wGrad, 4 * RC_NPARAM, hipMemcpyDeviceToHost, rc->stream), "read");
    if(tableGrad)
        g = g && ok(hipMemcpyAsync(tableGrad, rc->net.tabGrad, 4 * tf, hipMemcpyDeviceToHost, rc->stream), "read");
    return g && ok(hipStreamSynchronize(rc->stream), "read");
}

extern "C" bool DNX_rc_train_samples(DNXradianceCache* rc, const float* samples, uint32_t n, bool step)
{
    if(!rc || !samples || n == 0 || (n & 15u)) return false;
    if(!ensure_train(rc, n)) return false;
    if(dn_rc_train(rc->net, samples, n, rc->acts, rc->stream, rc->side.stream ? &rc->side : nullptr) != 0)
        return false;
    return !step || adam_step(rc);
}

extern "C" bool DNX_rc_predict_samples(DNXradianceCache* rc, const float* samples, uint32_t n, float* out)
{
    if(!rc || !samples || !out || n == 0 || (n & 15u)) return false;
    if(!ensure_query(rc, n)) return false;
    const DNrcNet inf = infer_net(rc);
    return dn_rc_encode(inf, samples, n, rc->dQueryX0, rc->stream) == 0 &&
           dn_rc_forward(inf, rc->dQueryX0, samples, n, out, rc->stream) == 0;
}

extern "C" bool DNX_rc_predict_samples_noalloc(DNXradianceCache* rc, const float* samples, uint32_t n, float* out)
{
    if(!rc || !samples || !out || n == 0 || (n & 15u)) return false;
    if(n > DNX_RC_NOALLOC_MAX_SAMPLES)
    {
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR,
               "radiance cache: %u samples exceed the allocation-free prediction's maximum of %u", n,
               DNX_RC_NOALLOC_MAX_SAMPLES);
        return false;
    }
    const int e = dn_rc_predict_noalloc(infer_net(rc), rc->device, samples, n, out, rc->stream);
    if(e == (int)hipErrorNotInitialized)
        dn_msg(DN_MESSAGE_GPU_MEMORY, DN_MESSAGE_ERROR,
               "radiance cache: no scratch array on device %d (DN_init was not called)", rc->device);
    else if(e)
        dn_msg(DN_MESSAGE_SHADER, DN_MESSAGE_ERROR, "radiance cache: allocation-free prediction failed (%d)", e);
    return e == 0;
}

extern "C" bool DNX_rc_train_frame(DNXradianceCache* rc, float time, int steps)
{
    if(!rc) return false;
    DNlightParams p;
    uint32_t maxReq = 0;
    if(!prepare_lighting(rc->vol, 1, 1000, time, p, maxReq)) return false;
    const uint32_t n = rc->cfg.samplesPerFrame;
    if(!ensure_train(rc, n)) return false;
    if(steps <= 0) return true;
    // Every step trains on samples of its own, drawn with the seed of its
    // frame counter, which advances per step: seed = base + frame * stride.
    // The samples of up to 16 steps are drawn in one launch (the draws of one
    // step, 2^17 lanes, fill two waves per SIMD; the training kernels of a step
    // do not read the volume, so drawing ahead changes nothing).
    const uint32_t group = steps < 16 ? (uint32_t)steps : 16u;
    if(!ensure_samples(rc, (size_t)group * n)) return false;
    for(int s0 = 0; s0 < steps; s0 += (int)group)
    {
        const uint32_t m = (uint32_t)(steps - s0) < group ? (uint32_t)(steps - s0) : group;
        const uint32_t seed = rc->cfg.seed * 0x9E3779B1u + rc->frame * 0x85EBCA77u + 0x632BE5ABu;
        if(dn_rc_samples(p, maxReq, n, m, seed, 0x85EBCA77u, rc->dSamples, rc->stream) != 0) return false;
        for(uint32_t k = 0; k < m; k++)
        {
            rc->frame++;
            const float* smp = rc->dSamples + (size_t)k * n * RC_SAMPLE;
            rc->lastSamplesPtr = smp;
            rc->lastSamples = n;
            if(dn_rc_train(rc->net, smp, n, rc->acts, rc->stream, rc->side.stream ? &rc->side : nullptr) != 0 ||
               !adam_step(rc))
                return false;
        }
    }
    return true;
}

extern "C" bool DNX_rc_predict_requests(DNXradianceCache* rc, float* out, size_t requestCapacity)
{
    if(!rc || !out) return false;
    DNlightParams p;
    uint32_t maxReq = 0;
    if(!prepare_lighting(rc->vol, 1, 1000, rc->vol->lastTime, p, maxReq)) return false;
    // the query rows' exact direct terms, traced ahead in direction-major waves
    // (bit-identical to tracing them per voxel; without them the gather does)
    trace_direct_terms(rc->vol, p, maxReq);
    size_t slots = (size_t)requestCapacity * 32u;
    if(slots == 0) return true;
    if(slots > 0xFFFFFFF0u || !ensure_query(rc, slots)) return false;
    const DNrcNet inf = infer_net(rc);
    return dn_rc_gather(p, (uint32_t)slots, rc->dQuery, rc->stream) == 0 &&
           dn_rc_encode(inf, rc->dQuery, (uint32_t)slots, rc->dQueryX0, rc->stream) == 0 &&
           dn_rc_forward(inf, rc->dQueryX0, rc->dQuery, (uint32_t)slots, out, rc->stream) == 0;
}

extern "C" bool DNX_rc_reference(DNXradianceCache* rc, double* acc, size_t requestCapacity, uint32_t samplesPerVoxel,
                                 uint32_t seed)
{
    if(!rc || !acc) return false;
    DNlightParams p;
    uint32_t maxReq = 0;
    if(!prepare_lighting(rc->vol, 1, 1000, rc->vol->lastTime, p, maxReq)) return false;
    const size_t slots = requestCapacity * 32u;
    if(slots > 0x7FFFFFF0u / 16u) return false;
    return dn_rc_reference(p, (uint32_t)slots, samplesPerVoxel, seed, acc, rc->stream) == 0;
}

extern "C" const float* DNX_rc_last_samples(DNXradianceCache* rc, uint32_t* n)
{
    if(!rc) return nullptr;
    if(n) *n = rc->lastSamples;
    return rc->lastSamplesPtr;
}

extern "C" bool DNX_accumulate_f64(DNvolume* vol, const float* src, double* dst, size_t n)
{
    if(!vol || !src || !dst) return false;
    return dn_accumulate_f64(src, dst, n, (hipStream_t)DNX_get_stream(vol)) == 0;
}

// mbrec_mex.hip -- the four MEX drop-ins over host FP64 buffers.
//
// Each entry point has the argument order and layouts of its reference MEX
// file.  Tables (travel times, SIR indices, SIR table, windows) are uploaded
// once and cached, keyed by their dimensions and a 64-bit hash of their full
// contents; images and sinograms are converted to FP32 planar layout, the
// device operator runs, and the result is widened back to FP64 in the MEX
// layout.  The per-call host<->device copies make this path PCIe-bound; the
// device API (mbrec_*_dev) is the one that runs at HBM speed.
#include "mbrec_internal.h"

#include <cstring>
#include <mutex>
#include <vector>
#include <list>

namespace mbrec {

static uint64_t hash_bytes(const void *data, size_t nbytes, uint64_t seed)
{
    // four interleaved multiply-xorshift lanes over 8-byte words (not cryptographic)
    const uint64_t M = 0x9E3779B97F4A7C15ull;
    uint64_t h[4] = {seed ^ 0x243F6A8885A308D3ull, seed ^ 0x13198A2E03707344ull,
                     seed ^ 0xA4093822299F31D0ull, seed ^ 0x082EFA98EC4E6C89ull};
    const uint8_t *p = (const uint8_t *)data;
    const size_t nw = nbytes / 8;
    size_t i = 0;
    for (; i + 4 <= nw; i += 4)
        for (int l = 0; l < 4; ++l) {
            uint64_t w;
            memcpy(&w, p + 8 * (i + l), 8);
            h[l] = (h[l] ^ w) * M;
            h[l] ^= h[l] >> 29;
        }
    for (; i < nw; ++i) {
        uint64_t w;
        memcpy(&w, p + 8 * i, 8);
        h[0] = (h[0] ^ w) * M;
        h[0] ^= h[0] >> 29;
    }
    uint64_t tail = 0;
    memcpy(&tail, p + 8 * nw, nbytes - 8 * nw);
    h[1] = (h[1] ^ tail ^ nbytes) * M;
    uint64_t r = h[0] ^ (h[1] * 3) ^ (h[2] * 5) ^ (h[3] * 7);
    r ^= r >> 31;
    return r * M;
}

struct CacheEntry {
    uint64_t key;
    int P, D, nT;
    double R;
    mbrec_plan plan;
};

static std::mutex g_cache_mu;
static std::list<CacheEntry> g_cache;   // most recently used first
static const size_t kCacheMax = 4;

static int get_plan(int P, int D, int nT, double R, const double *tt, const int32_t *sidx,
                    const double *sir, int Ls, int Nsir, const int32_t *se, mbrec_plan *out)
{
    uint64_t key = hash_bytes(tt, sizeof(double) * (size_t)P * D, 1469598103934665603ull);
    if (sidx) {
        key = hash_bytes(sidx, sizeof(int32_t) * (size_t)P * D, key);
        key = hash_bytes(sir, sizeof(double) * (size_t)Ls * Nsir, key ^ (uint64_t)Ls);
        key = hash_bytes(se, sizeof(int32_t) * 2 * (size_t)Nsir, key ^ (uint64_t)Nsir);
    }
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
        if (it->key == key && it->P == P && it->D == D && it->nT == nT && it->R == R) {
            g_cache.splice(g_cache.begin(), g_cache, it);
            *out = it->plan;
            return MBREC_OK;
        }
    mbrec_plan plan = nullptr;
    int rc = mbrec_plan_create(&plan, P, 1, D, nT, R, tt);
    if (rc) return rc;
    if (sidx && (rc = mbrec_plan_set_sir(plan, sidx, sir, Ls, Nsir, se))) {
        mbrec_plan_destroy(plan);
        return rc;
    }
    g_cache.push_front(CacheEntry{key, P, D, nT, R, plan});
    while (g_cache.size() > kCacheMax) {
        mbrec_plan_destroy(g_cache.back().plan);
        g_cache.pop_back();
    }
    *out = plan;
    return MBREC_OK;
}

struct DevBuf {
    float *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

static int run_forward(int W, int P, int D, int nT, double R, const double *image, const double *tt,
                       const int32_t *sidx, const double *sir, int Ls, int Nsir, const int32_t *se,
                       double *out)
{
    if (W <= 0 || P <= 0 || D <= 0 || nT <= 0 || !image || !tt || !out)
        return set_error(MBREC_ERR_ARG, "forwardMex:args", "invalid dimensions or null buffer");
    int rc = mbrec_init(0);
    if (rc) return rc;
    mbrec_plan plan;
    if ((rc = get_plan(P, D, nT, R, tt, sidx, sir, Ls, Nsir, se, &plan))) return rc;
    // image [W x P] -> planar [W][P]
    std::vector<float> img((size_t)W * P);
    for (int x = 0; x < P; ++x)
        for (int w = 0; w < W; ++w) img[(size_t)w * P + x] = (float)image[(size_t)x * W + w];
    DevBuf di, ds;
    const size_t nS = (size_t)W * D * nT;
    MBREC_HIP(hipMalloc((void **)&di.p, img.size() * sizeof(float)));
    MBREC_HIP(hipMalloc((void **)&ds.p, nS * sizeof(float)));
    MBREC_HIP(hipMemcpy(di.p, img.data(), img.size() * sizeof(float), hipMemcpyHostToDevice));
    if ((rc = op_forward(&plan->p, di.p, ds.p, W, false, nullptr))) return rc;
    std::vector<float> sino(nS);
    MBREC_HIP(hipMemcpy(sino.data(), ds.p, nS * sizeof(float), hipMemcpyDeviceToHost));
    // planar [W][D][nT] -> MEX [W x nT x D]
    for (int w = 0; w < W; ++w)
        for (int d = 0; d < D; ++d)
            for (int t = 0; t < nT; ++t)
                out[(size_t)d * W * nT + (size_t)t * W + w] = (double)sino[((size_t)w * D + d) * nT + t];
    return MBREC_OK;
}

static int run_transpose(int W, int nT, int D, int P, double R, const double *sig, const double *tt,
                         const int32_t *sidx, const double *sir, int Ls, int Nsir, const int32_t *se,
                         double *out)
{
    if (W <= 0 || P <= 0 || D <= 0 || nT <= 0 || !sig || !tt || !out)
        return set_error(MBREC_ERR_ARG, "forwardMex:args", "invalid dimensions or null buffer");
    int rc = mbrec_init(0);
    if (rc) return rc;
    mbrec_plan plan;
    if ((rc = get_plan(P, D, nT, R, tt, sidx, sir, Ls, Nsir, se, &plan))) return rc;
    const size_t nS = (size_t)W * D * nT;
    std::vector<float> sino(nS);
    for (int w = 0; w < W; ++w)
        for (int d = 0; d < D; ++d)
            for (int t = 0; t < nT; ++t)
                sino[((size_t)w * D + d) * nT + t] = (float)sig[(size_t)d * W * nT + (size_t)t * W + w];
    DevBuf di, ds;
    MBREC_HIP(hipMalloc((void **)&di.p, (size_t)W * P * sizeof(float)));
    MBREC_HIP(hipMalloc((void **)&ds.p, nS * sizeof(float)));
    MBREC_HIP(hipMemcpy(ds.p, sino.data(), nS * sizeof(float), hipMemcpyHostToDevice));
    if ((rc = op_transpose(&plan->p, ds.p, di.p, W, false, nullptr))) return rc;
    std::vector<float> img((size_t)W * P);
    MBREC_HIP(hipMemcpy(img.data(), di.p, img.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int x = 0; x < P; ++x)
        for (int w = 0; w < W; ++w) out[(size_t)x * W + w] = (double)img[(size_t)w * P + x];
    return MBREC_OK;
}

}  // namespace mbrec

using namespace mbrec;

extern "C" {

int mbrec_forwardMex(const double *image, int W, int P, const int32_t *sirIndices,
                     const double *travelTimes, int D, const double *sir, int Ls, int Nsir,
                     int32_t numTimesteps, const int32_t *se, double R, double *out)
{
    if (!sirIndices || !sir || !se)
        return set_error(MBREC_ERR_ARG, "forwardMex:args", "null SIR table");
    return run_forward(W, P, D, numTimesteps, R, image, travelTimes, sirIndices, sir, Ls, Nsir, se, out);
}

int mbrec_transposeMex(const double *sigMat, int W, int nT, int D, const int32_t *sirIndices,
                       const double *travelTimes, int P, const double *sir, int Ls, int Nsir,
                       const int32_t *se, double R, double *out)
{
    if (!sirIndices || !sir || !se)
        return set_error(MBREC_ERR_ARG, "forwardMex:args", "null SIR table");
    return run_transpose(W, nT, D, P, R, sigMat, travelTimes, sirIndices, sir, Ls, Nsir, se, out);
}

int mbrec_forwardDelayAndSumMex(const double *image, int W, int P, const double *travelTimes, int D,
                                int32_t numTimesteps, double R, double *out)
{
    return run_forward(W, P, D, numTimesteps, R, image, travelTimes, nullptr, nullptr, 0, 0, nullptr, out);
}

int mbrec_transposeDelayAndSumMex(const double *sigMat, int W, int nT, int D, const double *travelTimes,
                                  int P, double R, double *out)
{
    return run_transpose(W, nT, D, P, R, sigMat, travelTimes, nullptr, nullptr, 0, 0, nullptr, out);
}

// Reference messages, by MEX name and argument position (forwardMex.cpp:30-50,
// transposeMex.cpp:30-50, forwardDelayAndSumMex.cpp:27-38,
// transposeDelayAndSumMex.cpp:27-38).  All four files use the "forwardMex:" prefix.
int mbrec_mex_type_error(const char *mex_name, int arg, int expect_int32)
{
    static const char *fwd[] = {"image", "sirIndices", "travelTimes", "sir", "numTimesteps",
                                "sirTimestepsStartAndEnd", "circularInterpolationRadiusInSamples"};
    static const char *adj[] = {"sigMat", "sirIndices", "travelTimes", "sir", "sirTimestepsStartAndEnd",
                                "circularInterpolationRadiusInSamples"};
    static const char *fdas[] = {"image", "travelTimes", "numTimesteps", "circularInterpolationRadiusInSamples"};
    static const char *adas[] = {"sigMat", "travelTimes", "circularInterpolationRadiusInSamples"};
    const char *const *names = fwd;
    int n = 7;
    if (mex_name && !strcmp(mex_name, "transposeMex")) { names = adj; n = 6; }
    else if (mex_name && !strcmp(mex_name, "forwardDelayAndSumMex")) { names = fdas; n = 4; }
    else if (mex_name && !strcmp(mex_name, "transposeDelayAndSumMex")) { names = adas; n = 3; }
    const char *what = (arg >= 0 && arg < n) ? names[arg] : "argument";
    if (expect_int32)
        return set_error(MBREC_ERR_TYPE, "forwardMex:notInt32", "%s must be type int32.", what);
    return set_error(MBREC_ERR_TYPE, "forwardMex:notDouble", "%s must be type double.", what);
}

void mbrec_mex_cache_clear(void)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto &e : g_cache) mbrec_plan_destroy(e.plan);
    g_cache.clear();
}

}  // extern "C"

// mbrec_mex.hip -- the four MEX drop-ins over host FP64 buffers.
//
// Each entry point has the argument order and layouts of its reference MEX
// file.  Tables (travel times, SIR indices, SIR table, windows) are uploaded
// once per table and cached.  A cache entry is found by the tables' host
// addresses, dimensions and R plus a fingerprint of each table: a hash of
// kFpSamples evenly strided elements and of the first and last kFpEdge elements
// (about 0.3 MB read per call whatever the table size, instead of the 328 + 164 MB
// a full hash reads at 400^2).  MATLAB gives a modified array a new data address
// unless it is edited in place while unshared; such an in-place edit of a cached
// table is seen only where it touches a sampled element, so callers that edit a
// table in place drop the cache (mbrec_mex_cache_clear, "clear mex").
//
// The image or sinogram is copied to the device as FP64 in the MEX layout and
// permuted / narrowed to the device's planar FP32 layout there (and back), in
// staging buffers that live with the cache entry and grow only when a call needs
// more: after the first call with a given size a call allocates nothing.
#include "mbrec_internal.h"

#include <cstring>
#include <list>
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

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

static const size_t kFpSamples = 32768;   // strided elements per table
static const size_t kFpEdge = 1024;       // leading and trailing elements per table

// fingerprint of an array of n elements of esize bytes: its address, size and a hash of
// the strided sample and both ends
static uint64_t fingerprint(const void *data, size_t n, size_t esize, uint64_t seed)
{
    const uint8_t *p = (const uint8_t *)data;
    uint64_t h = hash_bytes(&data, sizeof(data), seed ^ (uint64_t)n * 0x100000001B3ull);
    const size_t edge = std::min(n, kFpEdge);
    h = hash_bytes(p, edge * esize, h);
    h = hash_bytes(p + (n - edge) * esize, edge * esize, h);
    if (n > 2 * kFpEdge) {
        const size_t stride = std::max<size_t>(1, n / kFpSamples);
        std::vector<uint8_t> buf;
        buf.reserve((n / stride + 1) * esize);
        for (size_t i = 0; i < n; i += stride) buf.insert(buf.end(), p + i * esize, p + (i + 1) * esize);
        h = hash_bytes(buf.data(), buf.size(), h);
    }
    return h;
}

struct CacheEntry {
    uint64_t key;
    int P, D, nT;
    double R;
    mbrec_plan plan;
    double *d_host_layout = nullptr;   // FP64 image / sinogram in the MEX layout
    float *d_img = nullptr;            // planar FP32 [W][P]
    float *d_sino = nullptr;           // planar FP32 [W][D][nT]
    int wcap = 0;                      // staging capacity in wavelengths
};

static void free_entry(CacheEntry &e)
{
    if (e.d_host_layout) (void)hipFree(e.d_host_layout);
    if (e.d_img) (void)hipFree(e.d_img);
    if (e.d_sino) (void)hipFree(e.d_sino);
    e.d_host_layout = nullptr;
    e.d_img = e.d_sino = nullptr;
    e.wcap = 0;
    mbrec_plan_destroy(e.plan);
}

static std::mutex g_cache_mu;
static std::list<CacheEntry> g_cache;   // most recently used first
static const size_t kCacheMax = 4;

// The MEX calls get P pixels without the grid shape.  The plan's device pixel order (16 x 16
// tiles) only pays off with the true column length Nz, so it is read off the travel times
// of the first detector: along a column (x = iz + Nz ix) they change smoothly, and they jump
// where one column ends and the next begins.  Nz = the common spacing of the jumps, accepted
// only if it divides P, every jump sits at a multiple of it, and the step between columns
// is smooth too; otherwise one column of P pixels (any factorisation gives the same result;
// the shape only decides which kernels are fast).
static int infer_column_length(const double *tt, int P)
{
    if (P < 64) return P;
    std::vector<double> dv(P - 1);
    for (int x = 0; x + 1 < P; ++x) dv[x] = std::fabs(tt[x + 1] - tt[x]);
    std::vector<double> srt(dv);
    std::nth_element(srt.begin(), srt.begin() + srt.size() / 2, srt.end());
    const double med = srt[srt.size() / 2];
    if (!(med > 0.0) || !std::isfinite(med)) return P;
    int first = -1;
    for (int x = 0; x + 1 < P; ++x)
        if (dv[x] > 8.0 * med) { first = x; break; }
    if (first < 0) return P;
    const int Nz = first + 1;
    if (Nz < 8 || P % Nz != 0 || P / Nz < 2) return P;
    double across = 0.0;   // mean step between neighbouring columns vs along a column
    for (int x = 0; x + 1 < P; ++x) {
        const bool boundary = (x + 1) % Nz == 0;
        if (!boundary && dv[x] > 8.0 * med) return P;   // a jump inside a column
        if (x + Nz < P) across += std::fabs(tt[x + Nz] - tt[x]);
    }
    across /= (double)(P - Nz);
    return across < 8.0 * med ? Nz : P;
}

// the cache entry of these tables, with staging for W wavelengths; the returned entry
// stays valid while g_cache_mu is held by the caller
static int get_entry(int P, int D, int nT, double R, const double *tt, const int32_t *sidx, const double *sir,
                     int Ls, int Nsir, const int32_t *se, int W, CacheEntry **out)
{
    uint64_t key = fingerprint(tt, (size_t)P * D, sizeof(double), 1469598103934665603ull);
    if (sidx) {
        key = fingerprint(sidx, (size_t)P * D, sizeof(int32_t), key);
        key = fingerprint(sir, (size_t)Ls * Nsir, sizeof(double), key ^ (uint64_t)Ls);
        key = fingerprint(se, 2 * (size_t)Nsir, sizeof(int32_t), key ^ (uint64_t)Nsir);
    }
    CacheEntry *e = nullptr;
    for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
        if (it->key == key && it->P == P && it->D == D && it->nT == nT && it->R == R) {
            g_cache.splice(g_cache.begin(), g_cache, it);
            e = &g_cache.front();
            break;
        }
    if (!e) {
        mbrec_plan plan = nullptr;
        const int Nz = infer_column_length(tt, P);
        int rc = mbrec_plan_create(&plan, Nz, P / Nz, D, nT, R, tt);
        if (rc) return rc;
        if (sidx && (rc = mbrec_plan_set_sir(plan, sidx, sir, Ls, Nsir, se))) {
            mbrec_plan_destroy(plan);
            return rc;
        }
        g_cache.push_front(CacheEntry{key, P, D, nT, R, plan});
        while (g_cache.size() > kCacheMax) {
            free_entry(g_cache.back());
            g_cache.pop_back();
        }
        e = &g_cache.front();
    }
    if (W > e->wcap) {
        if (e->d_host_layout) {
            MBREC_HIP(hipDeviceSynchronize());
            (void)hipFree(e->d_host_layout);
            (void)hipFree(e->d_img);
            (void)hipFree(e->d_sino);
            e->d_host_layout = nullptr;
            e->d_img = e->d_sino = nullptr;
            e->wcap = 0;
        }
        const size_t nimg = (size_t)W * P, nsino = (size_t)W * D * nT;
        MBREC_HIP(hipMalloc((void **)&e->d_host_layout, std::max(nimg, nsino) * sizeof(double)));
        MBREC_HIP(hipMalloc((void **)&e->d_img, nimg * sizeof(float)));
        MBREC_HIP(hipMalloc((void **)&e->d_sino, nsino * sizeof(float)));
        e->wcap = W;
    }
    *out = e;
    return MBREC_OK;
}

// [P x W] FP64 (wavelength fastest) <-> planar FP32 [W][P]
__global__ void k_mex_img_in(const double *__restrict__ a, float *__restrict__ b, int W, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t x = i / W, w = i - x * W;
        b[w * (n / W) + x] = (float)a[i];
    }
}

__global__ void k_mex_img_out(const float *__restrict__ b, double *__restrict__ a, int W, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t x = i / W, w = i - x * W;
        a[i] = (double)b[w * (n / W) + x];
    }
}

// [W x nT x D] FP64 (wavelength fastest) <-> planar FP32 [W][D][nT]
__global__ void k_mex_sino_in(const double *__restrict__ a, float *__restrict__ b, int W, int nT, int D)
{
    const size_t n = (size_t)W * nT * D;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t w = i % W, r = i / W, t = r % nT, d = r / nT;
        b[(w * D + d) * nT + t] = (float)a[i];
    }
}

__global__ void k_mex_sino_out(const float *__restrict__ b, double *__restrict__ a, int W, int nT, int D)
{
    const size_t n = (size_t)W * nT * D;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t w = i % W, r = i / W, t = r % nT, d = r / nT;
        a[i] = (double)b[(w * D + d) * nT + t];
    }
}

static dim3 grid_of(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 16384)); }

static int run_forward(int W, int P, int D, int nT, double R, const double *image, const double *tt,
                       const int32_t *sidx, const double *sir, int Ls, int Nsir, const int32_t *se,
                       double *out)
{
    if (W <= 0 || P <= 0 || D <= 0 || nT <= 0 || !image || !tt || !out)
        return set_error(MBREC_ERR_ARG, "forwardMex:args", "invalid dimensions or null buffer");
    int rc = mbrec_init(0);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    CacheEntry *e;
    if ((rc = get_entry(P, D, nT, R, tt, sidx, sir, Ls, Nsir, se, W, &e))) return rc;
    const size_t nimg = (size_t)W * P, nS = (size_t)W * D * nT;
    MBREC_HIP(hipMemcpy(e->d_host_layout, image, nimg * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mex_img_in, grid_of(nimg), dim3(256), 0, nullptr, (const double *)e->d_host_layout,
                       e->d_img, W, nimg);
    if ((rc = op_forward(&e->plan->p, e->d_img, e->d_sino, W, false, nullptr))) return rc;
    hipLaunchKernelGGL(k_mex_sino_out, grid_of(nS), dim3(256), 0, nullptr, (const float *)e->d_sino,
                       e->d_host_layout, W, nT, D);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipMemcpy(out, e->d_host_layout, nS * sizeof(double), hipMemcpyDeviceToHost));
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
    std::lock_guard<std::mutex> lk(g_cache_mu);
    CacheEntry *e;
    if ((rc = get_entry(P, D, nT, R, tt, sidx, sir, Ls, Nsir, se, W, &e))) return rc;
    const size_t nimg = (size_t)W * P, nS = (size_t)W * D * nT;
    MBREC_HIP(hipMemcpy(e->d_host_layout, sig, nS * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mex_sino_in, grid_of(nS), dim3(256), 0, nullptr, (const double *)e->d_host_layout,
                       e->d_sino, W, nT, D);
    if ((rc = op_transpose(&e->plan->p, e->d_sino, e->d_img, W, false, nullptr))) return rc;
    hipLaunchKernelGGL(k_mex_img_out, grid_of(nimg), dim3(256), 0, nullptr, (const float *)e->d_img,
                       e->d_host_layout, W, nimg);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipMemcpy(out, e->d_host_layout, nimg * sizeof(double), hipMemcpyDeviceToHost));
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
    if (!g_cache.empty()) (void)hipDeviceSynchronize();
    for (auto &e : g_cache) free_entry(e);
    g_cache.clear();
}

}  // extern "C"

// mbrec_backproj.hip -- secondary reconstruction paths on the device:
//   dmas_cf           (reconstruction_backprojection/dmas_cf.m:19-107)
//   backproject_waveeq (reconstruction_backprojection/backproject_waveeq.m:1-50)
//   bp_kernel          (reconstruction_backprojection/bp_kernel.m:1-12)
//   upsample_detectors_in_sinogram (load_and_preprocess/upsample_detectors_in_sinogram.m:1-13)
//
// These paths are FP64 in the reference and have no iterative error
// tolerance, so they stay FP64 here: gathers accumulate in FP64 and the dense
// contractions (K(idx,idx)*S, H*diag(w lambda)*H', the spline map) are FP64
// GEMMs on the matrix cores through rocBLAS.
#include "mbrec_internal.h"

#include <rocblas/rocblas.h>

#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

namespace mbrec {

// ---------------------------------------------------------------------------
// rocBLAS handle (one per process, stream set per call)
// ---------------------------------------------------------------------------
static rocblas_handle g_blas = nullptr;
static std::mutex g_blas_mu;

int blas_handle(hipStream_t s, rocblas_handle *out)
{
    std::lock_guard<std::mutex> lk(g_blas_mu);
    if (!g_blas && rocblas_create_handle(&g_blas) != rocblas_status_success)
        return set_error(MBREC_ERR_HIP, "mbrec:rocblas", "rocblas_create_handle failed");
    if (rocblas_set_stream(g_blas, s) != rocblas_status_success)
        return set_error(MBREC_ERR_HIP, "mbrec:rocblas", "rocblas_set_stream failed");
    *out = g_blas;
    return MBREC_OK;
}

static int dgemm(hipStream_t s, bool ta, bool tb, int m, int n, int k, double alpha, const double *A,
                 int lda, const double *B, int ldb, double beta, double *C, int ldc)
{
    rocblas_handle h;
    int rc = blas_handle(s, &h);
    if (rc) return rc;
    if (rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                      tb ? rocblas_operation_transpose : rocblas_operation_none, m, n, k, &alpha, A, lda, B,
                      ldb, &beta, C, ldc) != rocblas_status_success)
        return set_error(MBREC_ERR_HIP, "mbrec:rocblas", "rocblas_dgemm failed");
    return MBREC_OK;
}

__device__ __forceinline__ double matlab_round(double v)
{
    return v >= 0.0 ? floor(v + 0.5) : -floor(-v + 0.5);
}

// detector positions of the backprojection formulas: angles -(offset + p*cov/(D-1)) deg,
// position (R cos, R sin) (dmas_cf.m:95, backproject_waveeq.m:18)
static void detector_angles(int D, double coverage, std::vector<double> &ang)
{
    const double off = (180.0 - coverage) / 2.0;
    const double step = coverage / (D - 1);
    ang.resize(D);
    for (int p = 0; p < D; ++p) ang[p] = -(off + p * step) * M_PI / 180.0;
}

// nearest sample index (1-based) T = round((|x - q_p| / c - t1) * fs + 1)
__device__ __forceinline__ long long nearest_index(double x, double z, double qx, double qz, double c,
                                                   double t1, double fs)
{
    const double dx = fabs(x - qx), dz = fabs(z - qz);
    return (long long)matlab_round((sqrt(dx * dx + dz * dz) / c - t1) * fs + 1.0);
}

// ---------------------------------------------------------------------------
// DMAS-CF
// ---------------------------------------------------------------------------
__global__ void k_absmax(const float *__restrict__ x, size_t n, unsigned int *out)
{
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(x[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));   // non-negative floats order as uints
}

__global__ void k_zero_u32(unsigned int *p) { *p = 0u; }

// sn = s / max|s| and sh = sign(sn) sqrt(|sn|) in FP64 (dmas_cf.m:99-103)
__global__ void k_dmas_prep(const float *__restrict__ s, size_t n, const unsigned int *mx, double *__restrict__ sn,
                            double *__restrict__ sh)
{
    const double m = (double)__uint_as_float(*mx);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = (double)s[i] / m;
        sn[i] = v;
        sh[i] = (v > 0.0) ? sqrt(v) : ((v < 0.0) ? -sqrt(-v) : 0.0);
    }
}

__global__ __launch_bounds__(256) void k_dmas_gather(const double *__restrict__ xp, const double *__restrict__ zp,
                                                     int P, const double *__restrict__ sn,
                                                     const double *__restrict__ sh, int W, int D, int nT,
                                                     const double *__restrict__ qx, const double *__restrict__ qz,
                                                     double c, double t1, double fs, double *__restrict__ out,
                                                     int *err)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)P * W) return;
    const int w = (int)(idx / P), x = (int)(idx - (size_t)w * P);
    const double px = xp[x], pz = zp[x];
    double ssq = 0.0, ssum = 0.0, num = 0.0, den = 0.0;
    for (int p = 0; p < D; ++p) {
        const long long T = nearest_index(px, pz, qx[p], qz[p], c, t1, fs);
        if (T < 1 || T > nT) { atomicOr(err, 1); continue; }
        const size_t o = ((size_t)w * D + p) * nT + (size_t)(T - 1);
        const double a = sn[o], b = sh[o];
        ssq += b * b;
        ssum += b;
        num += a;
        den += a * a;
    }
    const double dmas = 0.5 * (ssum * ssum - ssq);
    const double cf = (num * num) / (1e-6 + D * den);
    const size_t WP = (size_t)W * P;
    out[idx] = dmas * cf;          // p0_rec_dmas_cf
    out[WP + idx] = dmas;          // p0_rec_dmas
    out[2 * WP + idx] = cf;        // cf
    out[3 * WP + idx] = num;       // p0_rec_das
}

// ---------------------------------------------------------------------------
// backproject_waveeq
// ---------------------------------------------------------------------------
// radial integrals: S = cumtrapz(t, s/c) ./ t along time, t_k = (k + 1 + cropped) dt.
// A workgroup per row: thread i takes 8 consecutive samples of a 2048-sample chunk, forms
// the trapezoid increments dt (s[k-1]/c + s[k]/c) / 2 and their running sum over its 8,
// a block scan of the thread totals gives each thread its offset (the chunk's total carried
// into the next chunk).  The partial sums are grouped by the scan instead of strictly left
// to right: FP64 rounding differences only.
__global__ __launch_bounds__(256) void k_cumtrapz(const double *__restrict__ s, double *__restrict__ out, int rows,
                                                  int nT, double c, double dt, int cropped)
{
    __shared__ double sc[256];
    const int r = blockIdx.x, t = threadIdx.x;
    if (r >= rows) return;   // the grid is one workgroup per row: uniform
    const double *a = s + (size_t)r * nT;
    double *o = out + (size_t)r * nT;
    double carry = 0.0;
    for (int k0 = 0; k0 < nT; k0 += 2048) {
        const int kb = k0 + 8 * t;
        double inc[8];
        double qp = (kb >= 1 && kb - 1 < nT) ? a[kb - 1] / c : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = kb + j;
            const double q = k < nT ? a[k] / c : 0.0;
            inc[j] = (k >= 1 && k < nT) ? dt * (qp + q) / 2.0 : 0.0;
            qp = q;
        }
#pragma unroll
        for (int j = 1; j < 8; ++j) inc[j] += inc[j - 1];
        sc[t] = inc[7];
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {   // inclusive scan of the thread totals
            const double v = t >= off ? sc[t - off] : 0.0;
            __syncthreads();
            sc[t] += v;
            __syncthreads();
        }
        const double base = carry + (t > 0 ? sc[t - 1] : 0.0);
        carry += sc[255];
        __syncthreads();   // sc is rewritten by the next chunk
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = kb + j;
            if (k < nT) o[k] = k == 0 ? 0.0 : (base + inc[j]) / ((k + 1 + cropped) * dt);
        }
    }
}

// 'same' 2-D convolution over (time, detector) per wavelength, then / sum(psf)
__global__ void k_conv_same(const double *__restrict__ s, double *__restrict__ out, int W, int D, int nT,
                            const double *__restrict__ psf, int pm, int pn, double inv_sum)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)W * D * nT) return;
    const int t = (int)(idx % nT);
    const int d = (int)((idx / nT) % D);
    const int w = (int)(idx / ((size_t)nT * D));
    const int ct = pm / 2, cd = pn / 2;
    double acc = 0.0;
    for (int b = 0; b < pn; ++b) {
        const int dd = d + cd - b;
        if (dd < 0 || dd >= D) continue;
        for (int a = 0; a < pm; ++a) {
            const int tt = t + ct - a;
            if (tt < 0 || tt >= nT) continue;
            acc += s[((size_t)w * D + dd) * nT + tt] * psf[(size_t)b * pm + a];
        }
    }
    out[idx] = acc * inv_sum;
}

// p0_x / p0_y accumulation (backproject_waveeq.m:35-47), FP64
__global__ __launch_bounds__(256) void k_bp_gather(const double *__restrict__ xp, const double *__restrict__ zp,
                                                   int P, const double *__restrict__ swl, int W, int D, int nT,
                                                   const double *__restrict__ qx, const double *__restrict__ qz,
                                                   const double *__restrict__ ca, const double *__restrict__ sa,
                                                   double c, double t1, double fs, double dtc,
                                                   double *__restrict__ ux, double *__restrict__ uy, int *err)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)P * W) return;
    const int w = (int)(idx / P), x = (int)(idx - (size_t)w * P);
    const double px = xp[x], pz = zp[x];
    double ax = 0.0, ay = 0.0;
    for (int p = 0; p < D; ++p) {
        const long long T = nearest_index(px, pz, qx[p], qz[p], c, t1, fs);
        if (T < 1 || T > nT) { atomicOr(err, 1); continue; }
        const double v = swl[((size_t)w * D + p) * nT + (size_t)(T - 1)];
        ax = ax + v * ca[p] * dtc;
        ay = ay + v * sa[p] * dtc;
    }
    ux[idx] = ax;
    uy[idx] = ay;
}

// fliplr(divergence(X, Y, U, V)) for meshgrid coordinates: x along columns (ix), y along rows
// (iz); central differences inside, one-sided at the edges (MATLAB gradient).
__global__ void k_divergence_flip(const double *__restrict__ ux, const double *__restrict__ uy,
                                  const double *__restrict__ xv, const double *__restrict__ zv, int Nz, int Nx,
                                  int W, double *__restrict__ out)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t P = (size_t)Nz * Nx;
    if (idx >= P * W) return;
    const int w = (int)(idx / P);
    const int pxl = (int)(idx - (size_t)w * P);
    const int ix = pxl / Nz, iz = pxl - ix * Nz;
    const double *U = ux + (size_t)w * P, *V = uy + (size_t)w * P;
    double dU, dV;
    if (Nx == 1) dU = 0.0;
    else if (ix == 0) dU = (U[(size_t)1 * Nz + iz] - U[iz]) / (xv[1] - xv[0]);
    else if (ix == Nx - 1) dU = (U[(size_t)ix * Nz + iz] - U[(size_t)(ix - 1) * Nz + iz]) / (xv[ix] - xv[ix - 1]);
    else dU = (U[(size_t)(ix + 1) * Nz + iz] - U[(size_t)(ix - 1) * Nz + iz]) / (xv[ix + 1] - xv[ix - 1]);
    if (Nz == 1) dV = 0.0;
    else if (iz == 0) dV = (V[(size_t)ix * Nz + 1] - V[(size_t)ix * Nz]) / (zv[1] - zv[0]);
    else if (iz == Nz - 1) dV = (V[(size_t)ix * Nz + iz] - V[(size_t)ix * Nz + iz - 1]) / (zv[iz] - zv[iz - 1]);
    else dV = (V[(size_t)ix * Nz + iz + 1] - V[(size_t)ix * Nz + iz - 1]) / (zv[iz + 1] - zv[iz - 1]);
    out[(size_t)w * P + (size_t)(Nx - 1 - ix) * Nz + iz] = dU + dV;   // fliplr: column ix -> Nx-1-ix
}

// ---------------------------------------------------------------------------
// bp_kernel: A = J0(t_j lambda_l) * (w_l lambda_l), B = Y0(t_j lambda_l); K = A B' - B A'
// ---------------------------------------------------------------------------
__global__ void k_hankel_parts(const double *__restrict__ t, const double *__restrict__ lam,
                               const double *__restrict__ wl, int Nt, double *__restrict__ A,
                               double *__restrict__ Bm)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)Nt * Nt) return;
    const int j = (int)(idx % Nt), l = (int)(idx / Nt);   // column-major [Nt x Nt]: row j, column l
    const double z = t[j] * lam[l];
    A[idx] = j0(z) * wl[l];
    Bm[idx] = y0(z);
}

__global__ void k_antisym(const double *__restrict__ M, double *__restrict__ K, int Nt)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)Nt * Nt) return;
    const int j = (int)(idx % Nt), k = (int)(idx / Nt);
    K[idx] = M[idx] - M[(size_t)j * Nt + k];
}

// scratch of one call, stream-ordered (no device-wide synchronisation on allocation or
// release, which a plain hipFree would add to every call)
struct DBuf {
    hipStream_t s;
    std::vector<void *> all;
    explicit DBuf(hipStream_t st) : s(st)
    {
        // keep up to 256 MB of released scratch in the device's default pool between calls,
        // so that a call's allocations do not map fresh memory each time (once per device)
        static std::atomic<unsigned long long> warmed{0};
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64 &&
            !(warmed.fetch_or(1ull << dev) & (1ull << dev))) {
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                uint64_t keep = (uint64_t)256 << 20;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
            (void)hipGetLastError();
        }
    }
    ~DBuf() { for (void *p : all) (void)hipFreeAsync(p, s); }
    template <typename T>
    int alloc(T **p, size_t n)
    {
        hipError_t e = hipMallocAsync((void **)p, (n ? n : 1) * sizeof(T), s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(MBREC_ERR_HIP, "mbrec:alloc", "allocation of %zu bytes failed", n * sizeof(T));
        }
        all.push_back(*p);
        return MBREC_OK;
    }
};

static inline int blocks_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 1u << 30); }

}  // namespace mbrec

using namespace mbrec;

extern "C" {

int mbrec_dmas_cf_dev(const double *xpix, const double *zpix, int P, const float *sig, int W, int D, int nT,
                      double c, double fs, int cropped, double coverage_deg, double radius, double *out,
                      mbrec_stream stream)
{
    if (!xpix || !zpix || !sig || !out || P <= 0 || W <= 0 || D <= 1 || nT <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "dmas_cf: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    std::vector<double> ang, hx(D), hz(D);
    detector_angles(D, coverage_deg, ang);
    for (int p = 0; p < D; ++p) { hx[p] = std::cos(ang[p]) * radius; hz[p] = std::sin(ang[p]) * radius; }
    DBuf b(s);
    double *qx, *qz, *sn, *sh;
    unsigned int *mx;
    int *err;
    const size_t n = (size_t)W * D * nT;
    int rc;
    if ((rc = b.alloc(&qx, D)) || (rc = b.alloc(&qz, D)) || (rc = b.alloc(&sn, n)) || (rc = b.alloc(&sh, n)) ||
        (rc = b.alloc(&mx, 1)) || (rc = b.alloc(&err, 1)))
        return rc;
    MBREC_HIP(hipMemcpyAsync(qx, hx.data(), sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(qz, hz.data(), sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, mx);
    hipLaunchKernelGGL(k_absmax, dim3(std::min(blocks_for(n), 2048)), dim3(256), 0, s, sig, n, mx);
    hipLaunchKernelGGL(k_dmas_prep, dim3(std::min(blocks_for(n), 4096)), dim3(256), 0, s, sig, n,
                       (const unsigned int *)mx, sn, sh);
    const double dt = 1.0 / fs, t1 = (1 + cropped) * dt;
    hipLaunchKernelGGL(k_dmas_gather, dim3(blocks_for((size_t)P * W)), dim3(256), 0, s, xpix, zpix, P,
                       (const double *)sn, (const double *)sh, W, D, nT, (const double *)qx, (const double *)qz,
                       c, t1, fs, out, err);
    int herr = 0;
    MBREC_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    MBREC_HIP(hipStreamSynchronize(s));
    if (herr)
        return set_error(MBREC_ERR_RANGE, "dmas_cf:indexOutOfRange",
                         "a pixel's arrival sample lies outside the sinogram (MATLAB raises an index error)");
    return MBREC_OK;
}

int mbrec_backproject_waveeq_dev(const double *xv, const double *zv, int Nz, int Nx, const double *sig, int W,
                                 int D, int nT, double c, double fs, int cropped, double coverage_deg,
                                 double radius, const double *psf_host, int pm, int pn, const double *K, int Nt,
                                 double *out, mbrec_stream stream)
{
    if (!xv || !zv || !sig || !K || !out || !psf_host || Nz <= 0 || Nx <= 0 || W <= 0 || D <= 1 || nT <= 0 ||
        pm <= 0 || pn <= 0 || Nt < nT + cropped)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "backproject_waveeq: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    const int P = Nz * Nx;
    std::vector<double> ang, hx(D), hz(D), hc(D), hs(D), hxp(P), hzp(P);
    detector_angles(D, coverage_deg, ang);
    for (int p = 0; p < D; ++p) {
        hc[p] = std::cos(ang[p]);
        hs[p] = std::sin(ang[p]);
        hx[p] = hc[p] * radius;
        hz[p] = hs[p] * radius;
    }
    double psum = 0.0;
    for (int i = 0; i < pm * pn; ++i) psum += psf_host[i];
    DBuf b(s);
    double *qx, *qz, *qc, *qs, *psf, *s1, *s2, *swl, *ux, *uy, *xp, *zp, *dxv, *dzv;
    int *err;
    const size_t n = (size_t)W * D * nT;
    int rc;
    if ((rc = b.alloc(&qx, D)) || (rc = b.alloc(&qz, D)) || (rc = b.alloc(&qc, D)) || (rc = b.alloc(&qs, D)) ||
        (rc = b.alloc(&psf, (size_t)pm * pn)) || (rc = b.alloc(&s1, n)) || (rc = b.alloc(&s2, n)) ||
        (rc = b.alloc(&swl, n)) || (rc = b.alloc(&ux, (size_t)P * W)) || (rc = b.alloc(&uy, (size_t)P * W)) ||
        (rc = b.alloc(&xp, P)) || (rc = b.alloc(&zp, P)) || (rc = b.alloc(&dxv, Nx)) || (rc = b.alloc(&dzv, Nz)) ||
        (rc = b.alloc(&err, 1)))
        return rc;
    // pixel coordinates from the grid vectors (meshgrid: x along columns, z along rows)
    for (int ix = 0; ix < Nx; ++ix)
        for (int iz = 0; iz < Nz; ++iz) {
            hxp[(size_t)ix * Nz + iz] = xv[ix];
            hzp[(size_t)ix * Nz + iz] = zv[iz];
        }
    MBREC_HIP(hipMemcpyAsync(dxv, xv, sizeof(double) * Nx, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(dzv, zv, sizeof(double) * Nz, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(qx, hx.data(), sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(qz, hz.data(), sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(qc, hc.data(), sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(qs, hs.data(), sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(psf, psf_host, sizeof(double) * pm * pn, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(xp, hxp.data(), sizeof(double) * P, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(zp, hzp.data(), sizeof(double) * P, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    const double dt = 1.0 / fs, t1 = (1 + cropped) * dt;
    hipLaunchKernelGGL(k_cumtrapz, dim3(W * D), dim3(256), 0, s, sig, s1, W * D, nT, c, dt,
                       cropped);
    hipLaunchKernelGGL(k_conv_same, dim3(blocks_for(n)), dim3(256), 0, s, (const double *)s1, s2, W, D, nT,
                       (const double *)psf, pm, pn, 1.0 / psum);
    // S_wl = K(idx, idx) * S * dt for all wavelengths at once: [nT x nT] x [nT x (D*W)]
    const double *Ksub = K + (size_t)cropped * Nt + cropped;
    if ((rc = dgemm(s, false, false, nT, D * W, nT, dt, Ksub, Nt, s2, nT, 0.0, swl, nT))) return rc;
    hipLaunchKernelGGL(k_bp_gather, dim3(blocks_for((size_t)P * W)), dim3(256), 0, s, (const double *)xp,
                       (const double *)zp, P, (const double *)swl, W, D, nT, (const double *)qx,
                       (const double *)qz, (const double *)qc, (const double *)qs, c, t1, fs, dt * c, ux, uy, err);
    hipLaunchKernelGGL(k_divergence_flip, dim3(blocks_for((size_t)P * W)), dim3(256), 0, s, (const double *)ux,
                       (const double *)uy, (const double *)dxv, (const double *)dzv, Nz, Nx, W, out);
    int herr = 0;
    MBREC_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    MBREC_HIP(hipStreamSynchronize(s));
    if (herr)
        return set_error(MBREC_ERR_RANGE, "backproject_waveeq:indexOutOfRange",
                         "a pixel's arrival sample lies outside the sinogram (MATLAB raises an index error)");
    return MBREC_OK;
}

int mbrec_bp_kernel_dev(const double *t_host, int Nt, double *K, mbrec_stream stream)
{
    if (!t_host || !K || Nt < 2) return set_error(MBREC_ERR_ARG, "mbrec:args", "bp_kernel: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    const double dl = M_PI / (t_host[Nt - 1] - t_host[0]);
    std::vector<double> lam(Nt), wl(Nt);
    for (int l = 0; l < Nt; ++l) {
        lam[l] = dl * (l + 1);
        const double wtr = (l == 0 || l == Nt - 1) ? 0.5 * dl : dl;   // trapz weights
        wl[l] = wtr * lam[l];
    }
    DBuf b(s);
    double *t, *dlam, *dwl, *A, *Bm, *M;
    const size_t nn = (size_t)Nt * Nt;
    int rc;
    if ((rc = b.alloc(&t, Nt)) || (rc = b.alloc(&dlam, Nt)) || (rc = b.alloc(&dwl, Nt)) || (rc = b.alloc(&A, nn)) ||
        (rc = b.alloc(&Bm, nn)) || (rc = b.alloc(&M, nn)))
        return rc;
    MBREC_HIP(hipMemcpyAsync(t, t_host, sizeof(double) * Nt, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(dlam, lam.data(), sizeof(double) * Nt, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(dwl, wl.data(), sizeof(double) * Nt, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_hankel_parts, dim3(blocks_for(nn)), dim3(256), 0, s, (const double *)t,
                       (const double *)dlam, (const double *)dwl, Nt, A, Bm);
    // M = A * B'  (A already carries the trapz weights times lambda)
    if ((rc = dgemm(s, false, true, Nt, Nt, Nt, 1.0, A, Nt, Bm, Nt, 0.0, M, Nt))) return rc;
    hipLaunchKernelGGL(k_antisym, dim3(blocks_for(nn)), dim3(256), 0, s, (const double *)M, K, Nt);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipStreamSynchronize(s));
    return MBREC_OK;
}

int mbrec_upsample_detectors_dev(const double *U, int Dup, int D, const double *sig, int nT, int nbatch,
                                 double *out, mbrec_stream stream)
{
    if (!U || !sig || !out || Dup <= 0 || D <= 0 || nT <= 0 || nbatch <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "upsample_detectors: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    // per batch item: out [nT x Dup] = sig [nT x D] * U' (U is [Dup x D] column-major)
    for (int b = 0; b < nbatch; ++b) {
        int rc = dgemm(s, false, true, nT, Dup, D, 1.0, sig + (size_t)b * D * nT, nT, U, Dup, 0.0,
                       out + (size_t)b * Dup * nT, nT);
        if (rc) return rc;
    }
    MBREC_HIP(hipStreamSynchronize(s));
    return MBREC_OK;
}

}  // extern "C"

// mbrec_preprocess.hip -- sinogram preprocessing on the device (load_and_preprocess/).
//
// filter_butter_zero_phase (filter_butter_zero_phase.m:1-53, FFT branch): demean along
// time when fmin > 0, then multiply the spectrum by |H|^2 of the cascade 8th-order
// Butterworth low-pass (fmax) x 2nd-order high-pass (fmin) sampled at the N = nT bins of the
// whole circle.  Multiplying the N-point DFT by a real even response is the circular
// convolution with k = ifft(|H|^2) (real, even), which is what runs here: an LDS-tiled FP64
// kernel, k computed once per (N, fs, fmin, fmax) by a direct DFT on the host.
//
// apply_butterworth_window_to_sinogram (apply_butterworth_window_to_sinogram.m:5-9): the
// magnitude |h| of the digital Butterworth band-pass [low high] / N (order n_butter) at the
// N points pi k / N of the upper half circle, used as a window along time around each
// column's mean: s = mean + (s - mean) |h|.
//
// interpolate_signals_of_broken_transducers (interpolate_signals_of_broken_transducers.m:6-18):
// interp2(..., 'spline') on a grid whose time axis is queried at its own knots reduces to
// the not-a-knot cubic spline across the working detectors, evaluated at every detector
// (extrapolating past the end ones): a [D x D] map applied per time sample.
//
// The Butterworth designs follow the textbook route MATLAB's butter takes: analog
// prototype poles exp(i pi (2k + n - 1) / (2n)), pre-warping u = 4 tan(pi Wn / 2) (fs = 2),
// low->low / low->high / low->band-pass transforms, bilinear transform.
#include "mbrec_internal.h"

#include <complex>
#include <mutex>
#include <vector>

namespace mbrec {

using cplx = std::complex<double>;

struct Zpk {
    std::vector<cplx> z, p;
    double k = 1.0;
};

// digital Butterworth: type 0 low-pass, 1 high-pass (w1 = cut-off), 2 band-pass [w1, w2];
// cut-offs normalised by the Nyquist rate
static Zpk butter_digital(int n, double w1, double w2, int type)
{
    const double pi = 3.14159265358979323846;
    std::vector<cplx> proto(n);
    for (int k = 1; k <= n; ++k) proto[k - 1] = std::exp(cplx(0.0, pi * (2.0 * k + n - 1) / (2.0 * n)));
    Zpk a;   // analog
    const double u1 = 4.0 * std::tan(pi * w1 / 2.0);
    if (type == 0) {
        for (const cplx &p : proto) a.p.push_back(u1 * p);
        a.k = std::pow(u1, n);
    } else if (type == 1) {
        for (const cplx &p : proto) a.p.push_back(u1 / p);
        a.z.assign(n, cplx(0.0, 0.0));
        a.k = 1.0;
    } else {
        const double u2 = 4.0 * std::tan(pi * w2 / 2.0);
        const double bw = u2 - u1, w0 = std::sqrt(u1 * u2);
        for (const cplx &p : proto) {
            const cplx h = p * bw / 2.0, r = std::sqrt(h * h - w0 * w0);
            a.p.push_back(h + r);
            a.p.push_back(h - r);
        }
        a.z.assign(n, cplx(0.0, 0.0));
        a.k = std::pow(bw, n);
    }
    // bilinear transform with fs = 2: s -> 4 (z - 1) / (z + 1)
    Zpk d;
    cplx num(1.0, 0.0), den(1.0, 0.0);
    for (const cplx &z : a.z) { d.z.push_back((4.0 + z) / (4.0 - z)); num *= 4.0 - z; }
    for (const cplx &p : a.p) { d.p.push_back((4.0 + p) / (4.0 - p)); den *= 4.0 - p; }
    d.k = a.k * std::real(num / den);
    while (d.z.size() < d.p.size()) d.z.push_back(cplx(-1.0, 0.0));   // zeros at Nyquist
    return d;
}

static double mag(const Zpk &f, double w)
{
    const cplx e = std::exp(cplx(0.0, w));
    cplx h(f.k, 0.0);
    for (const cplx &z : f.z) h *= e - z;
    for (const cplx &p : f.p) h /= e - p;
    return std::abs(h);
}

// kernel k = ifft(|H_lp H_hp|^2 at 2 pi f / N), cached per (N, fs, fmin, fmax)
struct FiltCache {
    int N = 0;
    double fs = 0, fmin = 0, fmax = 0;
    double *d_k = nullptr;
};
static std::mutex g_filt_mu;
static FiltCache g_filt;

static int filter_kernel(int N, double fs, double fmin, double fmax, const double **out)
{
    std::lock_guard<std::mutex> lk(g_filt_mu);
    if (g_filt.d_k && g_filt.N == N && g_filt.fs == fs && g_filt.fmin == fmin && g_filt.fmax == fmax) {
        *out = g_filt.d_k;
        return MBREC_OK;
    }
    const double pi = 3.14159265358979323846;
    std::vector<double> H2(N, 1.0);
    if (fmax > 0) {
        const Zpk f = butter_digital(8, fmax * 2.0 / fs, 0.0, 0);
        for (int i = 0; i < N; ++i) { const double m = mag(f, 2.0 * pi * i / N); H2[i] *= m * m; }
    }
    if (fmin > 0) {
        const Zpk f = butter_digital(2, fmin * 2.0 / fs, 0.0, 1);
        for (int i = 0; i < N; ++i) { const double m = mag(f, 2.0 * pi * i / N); H2[i] *= m * m; }
    }
    std::vector<double> c(N), k(N);
    for (int i = 0; i < N; ++i) c[i] = std::cos(2.0 * pi * i / N);
    for (int j = 0; j < N; ++j) {
        double acc = 0.0;
        long long idx = 0;
        for (int f = 0; f < N; ++f) {
            acc += H2[f] * c[idx];
            idx += j;
            if (idx >= N) idx -= N;
        }
        k[j] = acc / N;
    }
    if (g_filt.d_k) (void)hipFree(g_filt.d_k);
    g_filt.d_k = nullptr;
    MBREC_HIP(hipMalloc((void **)&g_filt.d_k, sizeof(double) * N));
    MBREC_HIP(hipMemcpy(g_filt.d_k, k.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    g_filt.N = N;
    g_filt.fs = fs;
    g_filt.fmin = fmin;
    g_filt.fmax = fmax;
    *out = g_filt.d_k;
    return MBREC_OK;
}

// per-column mean of [ncols][N] -> mean[ncols] (one block per column, FP64)
__global__ __launch_bounds__(256) void k_col_mean(const double *__restrict__ x, double *__restrict__ mean, int N)
{
    __shared__ double red[256];
    const double *c = x + (size_t)blockIdx.x * N;
    double s = 0.0;
    for (int i = threadIdx.x; i < N; i += 256) s += c[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) mean[blockIdx.x] = red[0] / N;
}

// y[c][t] = sum_m k[(t - m) mod N] (x[c][m] - mean[c]): block = 256 outputs t x kCB columns;
// per chunk of 256 m the 511 kernel values it needs and the kCB x 256 inputs go to LDS
constexpr int kCB = 8;
__global__ __launch_bounds__(256) void k_circ_conv(const double *__restrict__ x, const double *__restrict__ mean,
                                                   const double *__restrict__ kern, double *__restrict__ y, int N,
                                                   int ncols)
{
    __shared__ double ks[512];
    __shared__ double xs[256][kCB];
    const int t0 = blockIdx.x * 256, c0 = blockIdx.y * kCB;
    const int tl = threadIdx.x, t = t0 + tl;
    double acc[kCB];
#pragma unroll
    for (int c = 0; c < kCB; ++c) acc[c] = 0.0;
    for (int m0 = 0; m0 < N; m0 += 256) {
        __syncthreads();
        // ks[i] = k[(t0 - m0 - 255 + i) mod N], i in [0, 511)
        for (int i = tl; i < 511; i += 256) {
            long long j = (long long)t0 - m0 - 255 + i;
            j %= N;
            if (j < 0) j += N;
            ks[i] = kern[j];
        }
        {
            const int m = m0 + tl;
#pragma unroll
            for (int c = 0; c < kCB; ++c) {
                const int col = c0 + c;
                xs[tl][c] = (m < N && col < ncols) ? x[(size_t)col * N + m] - (mean ? mean[col] : 0.0) : 0.0;
            }
        }
        __syncthreads();
        const int mlim = min(256, N - m0);
        for (int jm = 0; jm < mlim; ++jm) {
            const double kv = ks[tl - jm + 255];
#pragma unroll
            for (int c = 0; c < kCB; ++c) acc[c] = fma(kv, xs[jm][c], acc[c]);
        }
    }
    if (t < N)
#pragma unroll
        for (int c = 0; c < kCB; ++c)
            if (c0 + c < ncols) y[(size_t)(c0 + c) * N + t] = acc[c];
}

// s = mean + (s - mean) * w[t]  (in place; [ncols][N])
__global__ __launch_bounds__(256) void k_window(double *__restrict__ s, const double *__restrict__ mean,
                                                const double *__restrict__ w, int N, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t c = i / N, t = i - c * N;
        const double mu = mean[c];
        s[i] = mu + (s[i] - mu) * w[t];
    }
}

// out[b][d][t] = sum_j U[d][j] x[b][j][t]: block = 256 time samples x one detector d
// (U row in LDS), j over the D input detectors
__global__ __launch_bounds__(256) void k_detector_map(const double *__restrict__ U, const double *__restrict__ x,
                                                      double *__restrict__ out, int D, int Dout, int N)
{
    extern __shared__ double urow[];
    const int d = blockIdx.y, b = blockIdx.z;
    for (int j = threadIdx.x; j < D; j += 256) urow[j] = U[(size_t)d * D + j];
    __syncthreads();
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= N) return;
    const double *xb = x + (size_t)b * D * N;
    double acc = 0.0;
    for (int j = 0; j < D; ++j) {
        const double u = urow[j];
        if (u != 0.0) acc = fma(u, xb[(size_t)j * N + t], acc);
    }
    out[((size_t)b * Dout + d) * N + t] = acc;
}

// not-a-knot cubic spline through (knots[j], e_i) evaluated at q: returns W [nq x nk] with
// W[q][i] the weight of value i (MATLAB spline / interp2 'spline' along one axis; nk >= 4;
// nk = 3: the parabola, nk = 2: the line, nk = 1: the constant)
static int spline_map(const std::vector<double> &kn, const std::vector<double> &q, std::vector<double> &W)
{
    const int n = (int)kn.size(), nq = (int)q.size();
    W.assign((size_t)nq * n, 0.0);
    if (n == 1) {
        for (int a = 0; a < nq; ++a) W[(size_t)a * n] = 1.0;
        return MBREC_OK;
    }
    for (int i = 0; i < n; ++i) {
        std::vector<double> y(n, 0.0);
        y[i] = 1.0;
        // second-derivative form: solve for M (curvatures) with not-a-knot end conditions
        std::vector<double> M(n, 0.0);
        if (n >= 4) {
            std::vector<double> h(n - 1);
            for (int j = 0; j + 1 < n; ++j) h[j] = kn[j + 1] - kn[j];
            // dense n x n system (n is a detector count): rows 1..n-2 continuity of y',
            // row 0 / n-1 continuity of y''' at the second / second-to-last knot
            std::vector<double> A((size_t)n * n, 0.0), r(n, 0.0);
            for (int j = 1; j + 1 < n; ++j) {
                A[(size_t)j * n + j - 1] = h[j - 1] / 6.0;
                A[(size_t)j * n + j] = (h[j - 1] + h[j]) / 3.0;
                A[(size_t)j * n + j + 1] = h[j] / 6.0;
                r[j] = (y[j + 1] - y[j]) / h[j] - (y[j] - y[j - 1]) / h[j - 1];
            }
            A[0] = 1.0 / h[0];
            A[1] = -1.0 / h[0] - 1.0 / h[1];
            A[2] = 1.0 / h[1];
            A[(size_t)(n - 1) * n + n - 3] = 1.0 / h[n - 3];
            A[(size_t)(n - 1) * n + n - 2] = -1.0 / h[n - 3] - 1.0 / h[n - 2];
            A[(size_t)(n - 1) * n + n - 1] = 1.0 / h[n - 2];
            // Gaussian elimination with partial pivoting
            std::vector<int> piv(n);
            for (int c = 0; c < n; ++c) {
                int best = c;
                for (int rr = c + 1; rr < n; ++rr)
                    if (std::fabs(A[(size_t)rr * n + c]) > std::fabs(A[(size_t)best * n + c])) best = rr;
                if (A[(size_t)best * n + c] == 0.0)
                    return set_error(MBREC_ERR_ARG, "mbrec:spline", "singular spline system");
                if (best != c) {
                    for (int cc = 0; cc < n; ++cc) std::swap(A[(size_t)c * n + cc], A[(size_t)best * n + cc]);
                    std::swap(r[c], r[best]);
                }
                for (int rr = c + 1; rr < n; ++rr) {
                    const double f = A[(size_t)rr * n + c] / A[(size_t)c * n + c];
                    if (f == 0.0) continue;
                    for (int cc = c; cc < n; ++cc) A[(size_t)rr * n + cc] -= f * A[(size_t)c * n + cc];
                    r[rr] -= f * r[c];
                }
            }
            for (int c = n - 1; c >= 0; --c) {
                double s = r[c];
                for (int cc = c + 1; cc < n; ++cc) s -= A[(size_t)c * n + cc] * M[cc];
                M[c] = s / A[(size_t)c * n + c];
            }
        }
        for (int a = 0; a < nq; ++a) {
            double v;
            if (n >= 4) {
                int j = 0;   // piece: the last knot interval whose left end is <= q (ends extended)
                while (j + 2 < n && q[a] >= kn[j + 1]) ++j;
                const double hj = kn[j + 1] - kn[j], A0 = (kn[j + 1] - q[a]) / hj, A1 = (q[a] - kn[j]) / hj;
                v = A0 * y[j] + A1 * y[j + 1] + ((A0 * A0 * A0 - A0) * M[j] + (A1 * A1 * A1 - A1) * M[j + 1]) * hj * hj / 6.0;
            } else if (n == 3) {   // interpolating parabola (Lagrange form)
                v = 0.0;
                for (int j = 0; j < 3; ++j) {
                    double l = y[j];
                    for (int m = 0; m < 3; ++m)
                        if (m != j) l *= (q[a] - kn[m]) / (kn[j] - kn[m]);
                    v += l;
                }
            } else {
                v = y[0] + (y[1] - y[0]) * (q[a] - kn[0]) / (kn[1] - kn[0]);
            }
            W[(size_t)a * n + i] = v;
        }
    }
    return MBREC_OK;
}

static dim3 grid1(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 16384)); }

}  // namespace mbrec

using namespace mbrec;

extern "C" {

int mbrec_filter_butter_zero_phase_dev(const double *sig, int nT, int ncols, double fs, double fmin, double fmax,
                                       double *out, mbrec_stream stream)
{
    if (!sig || !out || sig == out || nT <= 0 || ncols <= 0 || !(fs > 0) || fmin < 0 || fmax < 0 ||
        fmin * 2.0 >= fs || fmax * 2.0 >= fs)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "filter_butter_zero_phase: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    const double *kern;
    int rc = filter_kernel(nT, fs, fmin, fmax, &kern);
    if (rc) return rc;
    double *mean = nullptr;
    if (fmin > 0) {   // filter_butter_zero_phase.m:17-21: demean only with a high-pass edge
        MBREC_HIP(hipMallocAsync((void **)&mean, sizeof(double) * ncols, s));
        hipLaunchKernelGGL(k_col_mean, dim3(ncols), dim3(256), 0, s, sig, mean, nT);
    }
    hipLaunchKernelGGL(k_circ_conv, dim3((nT + 255) / 256, (ncols + kCB - 1) / kCB), dim3(256), 0, s, sig,
                       (const double *)mean, kern, out, nT, ncols);
    MBREC_HIP(hipGetLastError());
    if (mean) MBREC_HIP(hipFreeAsync(mean, s));
    MBREC_HIP(hipStreamSynchronize(s));
    return MBREC_OK;
}

int mbrec_apply_butterworth_window_dev(double *sig, int nT, int ncols, int n_butter, double low_cutoff,
                                       double high_cutoff, mbrec_stream stream)
{
    if (!sig || nT <= 0 || ncols <= 0 || n_butter <= 0 || !(low_cutoff > 0) || !(high_cutoff > low_cutoff) ||
        high_cutoff >= nT)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "apply_butterworth_window_to_sinogram: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    const double pi = 3.14159265358979323846;
    const Zpk f = butter_digital(n_butter, low_cutoff / nT, high_cutoff / nT, 2);
    std::vector<double> w(nT);
    for (int t = 0; t < nT; ++t) w[t] = mag(f, pi * t / nT);
    double *dw = nullptr, *mean = nullptr;
    MBREC_HIP(hipMallocAsync((void **)&dw, sizeof(double) * nT, s));
    MBREC_HIP(hipMallocAsync((void **)&mean, sizeof(double) * ncols, s));
    MBREC_HIP(hipMemcpyAsync(dw, w.data(), sizeof(double) * nT, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_col_mean, dim3(ncols), dim3(256), 0, s, (const double *)sig, mean, nT);
    hipLaunchKernelGGL(k_window, grid1((size_t)nT * ncols), dim3(256), 0, s, sig, (const double *)mean,
                       (const double *)dw, nT, (size_t)nT * ncols);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipFreeAsync(dw, s));
    MBREC_HIP(hipFreeAsync(mean, s));
    MBREC_HIP(hipStreamSynchronize(s));   // w is a host temporary
    return MBREC_OK;
}

int mbrec_interpolate_broken_transducers_dev(const double *sig, int D, int nT, int nbatch, const int32_t *broken,
                                             int nbroken, double *out, mbrec_stream stream)
{
    if (!sig || !out || sig == out || D <= 0 || nT <= 0 || nbatch <= 0 || nbroken < 0 || (nbroken && !broken))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "interpolate_signals_of_broken_transducers: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    std::vector<char> bad(D, 0);
    for (int i = 0; i < nbroken; ++i) {
        if (broken[i] < 1 || broken[i] > D)
            return set_error(MBREC_ERR_ARG, "mbrec:args", "broken transducer %d outside [1, %d]", broken[i], D);
        bad[broken[i] - 1] = 1;
    }
    std::vector<double> kn, q(D);
    std::vector<int> good;
    for (int d = 0; d < D; ++d) {
        q[d] = d + 1.0;
        if (!bad[d]) { kn.push_back(d + 1.0); good.push_back(d); }
    }
    if (kn.empty())
        return set_error(MBREC_ERR_ARG, "mbrec:args", "interpolate_signals_of_broken_transducers: no working transducer");
    std::vector<double> Wm;
    int rc = spline_map(kn, q, Wm);
    if (rc) return rc;
    std::vector<double> U((size_t)D * D, 0.0);   // [D out][D in], zero columns for broken inputs
    for (int d = 0; d < D; ++d)
        for (size_t i = 0; i < good.size(); ++i) U[(size_t)d * D + good[i]] = Wm[(size_t)d * good.size() + i];
    double *dU = nullptr;
    MBREC_HIP(hipMallocAsync((void **)&dU, sizeof(double) * U.size(), s));
    MBREC_HIP(hipMemcpyAsync(dU, U.data(), sizeof(double) * U.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_detector_map, dim3((nT + 255) / 256, D, nbatch), dim3(256), sizeof(double) * D, s,
                       (const double *)dU, sig, out, D, D, nT);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipFreeAsync(dU, s));
    MBREC_HIP(hipStreamSynchronize(s));   // U is a host temporary
    return MBREC_OK;
}

}  // extern "C"

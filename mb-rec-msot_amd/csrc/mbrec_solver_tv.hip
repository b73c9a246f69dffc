// mbrec_solver_tv.hip -- device rec_nn_with_TV_reg.
//
// rec_nn_with_TV_reg.m:20-53 -> nnls_sparsa_based_reg_matrix_less.m:24-86: the same
// outer splitting as the L2 loop, the inner problem solved by SpaRSA (published
// algorithm: separable step x+ = Psi(x - grad/alpha, tau/alpha), every step accepted,
// Barzilai-Borwein alpha, initial alpha 1; no fewer than MiniterA and, past MiniterA,
// no more than MaxiterA steps, stopping early on a relative objective change <= 1e-8;
// the reference passes MaxiterA = 2 and leaves MiniterA at SpaRSA's default of 5) with
//   Psi = isotropic-TV prox by split Bregman (SB_ITV2.m:20-40, stop at
//         ||u_prev - u|| / ||u|| <= 1e-3, shrinkage with d = 0 where |B u + b| = 0),
//   Phi = sum of isotropic TV with central differences (ITV.m:5-6).
// The inner linear solve of SB_ITV2, cgs(I + B'B, rhs, 1e-5, 100), is done exactly
// here: for the DiffOper B (backward differences, zero first row) B'B is the 2-D
// Neumann Laplacian, diagonalised by the orthonormal DCT-II basis, so
// (I + B'B)^-1 r = C ((C' R C) ./ (1 + l_i + l_j)) C' with l_k = 4 sin^2(pi k / 2N).
// The 1-D transforms are FP64 DCTs computed from one complex FFT of length N per two
// real rows (Makhoul's even/odd reordering; the two rows as real and imaginary part),
// the FFT a mixed-radix Stockham transform in LDS over N's prime factors (radix 4
// first), O(N^2 log N) per image for smooth N.  Three passes per solve: along z (the
// right-hand side g - B'(b - d) formed while loading), along x (forward, divide by the
// eigenvalues, inverse), along z (inverse).  The reference's cgs stops at a relative
// residual of 1e-5.
//
// Model-facing vectors are FP32, the TV prox is FP64, scalars are FP64.  The loops are
// host-driven (the split Bregman loop length depends on the data).
#include "mbrec_reduce.h"

#include <cmath>
#include <vector>

namespace mbrec {

// ---- TV prox kernels (images [B][Nx][Nz] column-major = z fastest) -------------
// g = (double) z
__global__ void k_tv_load(const float *__restrict__ z, double *__restrict__ g, double *__restrict__ u,
                          double *__restrict__ bz, double *__restrict__ bx, double *__restrict__ dz,
                          double *__restrict__ dx, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = z[i];
        g[i] = v;
        u[i] = v;
        bz[i] = bx[i] = dz[i] = dx[i] = 0.0;
    }
}

// ---- Neumann solve: DCTs by mixed-radix FFT in LDS -------------------------------
constexpr int kNmMaxFactors = 24;
constexpr int kNmLdsPoints = 2048;   // complex points per LDS buffer (two buffers: 64 KB)
constexpr int kNmMaxN = kNmLdsPoints;

struct NmFactors {
    int n;
    int p[kNmMaxFactors];
};

// position of sample i in Makhoul's reordering: even samples first, odd ones reversed
__device__ inline int nm_perm(int i, int N) { return (i & 1) ? N - 1 - (i >> 1) : (i >> 1); }

// Stockham FFT of C rows of N complex points (row c at a + c*N), unnormalised,
// e^{-2 pi i/N} (forward) or e^{+2 pi i/N} (inverse) kernel; tw[j] = e^{-2 pi i j/N}.
// Stage with radix p after Ns points per sub-transform: output (j, q), j < N/p, q < p,
// goes to (j / Ns) Ns p + j % Ns + q Ns and is sum_r a[j + r N/p] w^(r t),
// t = (j % Ns + q Ns) N / (Ns p).  Returns the buffer holding the result.
// one output of a radix-p stage: sum_r src[r m] w^(r t) (P > 0: p known at compile time)
template <int P>
__device__ __forceinline__ double2 nm_bfly(const double2 *src, int p, int m, int t, int N,
                                           const double2 *__restrict__ tw, bool inverse)
{
    double re = 0.0, im = 0.0;
    int idx = 0;
#pragma unroll
    for (int r = 0; r < (P ? P : p); ++r) {
        const double2 w = tw[idx];
        const double wi = inverse ? -w.y : w.y;
        const double2 x = src[r * m];
        re += x.x * w.x - x.y * wi;
        im += x.x * wi + x.y * w.x;
        idx += t;
        if (idx >= N) idx -= N;
    }
    return make_double2(re, im);
}

__device__ double2 *nm_fft(double2 *a, double2 *b, int N, int C, const NmFactors &f,
                           const double2 *__restrict__ tw, bool inverse)
{
    int Ns = 1;
    const float invN = 1.0f / (float)N;
    for (int s = 0; s < f.n; ++s) {
        const int p = f.p[s], m = N / p, step = N / (Ns * p);
        const float invm = 1.0f / (float)m, invNs = 1.0f / (float)Ns;   // the stage's divisions as reciprocals
        for (int o = threadIdx.x; o < C * N; o += blockDim.x) {
            int c, e, q, j, jn, k;
            divmod_small(o, N, invN, c, e);
            divmod_small(e, m, invm, q, j);
            divmod_small(j, Ns, invNs, jn, k);
            const int t = (k + q * Ns) * step;
            const double2 *src = a + c * N + j;
            double2 y;
            switch (p) {   // uniform: the radices nm_factors produces first
            case 2: y = nm_bfly<2>(src, p, m, t, N, tw, inverse); break;
            case 3: y = nm_bfly<3>(src, p, m, t, N, tw, inverse); break;
            case 4: y = nm_bfly<4>(src, p, m, t, N, tw, inverse); break;
            case 5: y = nm_bfly<5>(src, p, m, t, N, tw, inverse); break;
            case 7: y = nm_bfly<7>(src, p, m, t, N, tw, inverse); break;
            default: y = nm_bfly<0>(src, p, m, t, N, tw, inverse); break;
            }
            b[c * N + jn * Ns * p + k + q * Ns] = y;
        }
        __syncthreads();
        double2 *sw = a;
        a = b;
        b = sw;
        Ns *= p;
    }
    return a;
}

// Orthonormal DCT-II of the two real rows packed in Z = FFT(v1 + i v2) (v the reordered
// rows): X_a[k] = s_k Re(e^{-i pi k/2N} V_a[k]), V1 = (Z[k] + conj Z[N-k]) / 2,
// V2 = (Z[k] - conj Z[N-k]) / 2i; s_0 = sqrt(1/N), s_k = sqrt(2/N).
__device__ inline double2 nm_dct_pair(const double2 *Z, int k, int N, const double2 *__restrict__ qw)
{
    const double2 z = Z[k], zc = Z[k == 0 ? 0 : N - k];
    const double v1r = 0.5 * (z.x + zc.x), v1i = 0.5 * (z.y - zc.y);
    const double v2r = 0.5 * (z.y + zc.y), v2i = -0.5 * (z.x - zc.x);
    const double2 q = qw[k];
    const double s = k == 0 ? sqrt(1.0 / N) : sqrt(2.0 / N);
    return make_double2(s * (q.x * v1r - q.y * v1i), s * (q.x * v2r - q.y * v2i));
}

// Inverse (orthonormal DCT-III) input for the two real rows X1, X2, given X_a[k] and
// X_a[N-k] (0 for k = 0): Y = X / s, W = Y[k] - i Y[N-k], V = e^{i pi k/2N} W, and the
// packed spectrum V1 + i V2 whose inverse FFT / N is v1 + i v2.
__device__ inline double2 nm_idct_pair(double2 xk, double2 xr, int k, int N, const double2 *__restrict__ qw)
{
    const double is = k == 0 ? sqrt((double)N) : sqrt(0.5 * N);
    const double isr = sqrt(0.5 * N);
    const double y1 = xk.x * is, y1r = k == 0 ? 0.0 : xr.x * isr;
    const double y2 = xk.y * is, y2r = k == 0 ? 0.0 : xr.y * isr;
    const double2 q = qw[k];                   // conj(q) = e^{i pi k/2N}
    const double v1r = q.x * y1 - q.y * y1r, v1i = -q.x * y1r - q.y * y1;
    const double v2r = q.x * y2 - q.y * y2r, v2i = -q.x * y2r - q.y * y2;
    return make_double2(v1r - v2i, v1i + v2r);
}

struct NmArgs {
    int N, C;
    NmFactors f;
    const double2 *tw, *qw;
    const double *lam;
};

// Pass 1, along z: rows = pairs of image columns (ix, ix + 1), loaded as the right-hand
// side rhs = g - B'(b - d),
//   (B' v)(iz,ix) = vz(iz,ix)[iz>0] - vz(iz+1,ix)[iz+1<N] + vx(iz,ix)[ix>0] - vx(iz,ix+1)[ix+1<N],
// and written as their DCT-II along z to t.  Grid (ceil(ceil(N/2) / C), B).
__global__ __launch_bounds__(256) void k_nm_z_fwd(const double *__restrict__ g, const double *__restrict__ bz,
                                                  const double *__restrict__ bx, const double *__restrict__ dz,
                                                  const double *__restrict__ dx, double *__restrict__ t,
                                                  const int *__restrict__ active, NmArgs A)
{
    extern __shared__ double2 nm_sm[];
    const int b = blockIdx.y;
    if (!active[b]) return;
    const int N = A.N, C = A.C;
    const size_t o = (size_t)b * N * N;
    double2 *a0 = nm_sm, *a1 = nm_sm + C * N;
    const int col0 = blockIdx.x * 2 * C;
    auto rhs = [&](int iz, int ix) -> double {
        if (ix >= N) return 0.0;
        const size_t k = o + (size_t)ix * N + iz;
        double bt = 0.0;
        if (iz > 0) bt += bz[k] - dz[k];
        if (iz + 1 < N) bt -= bz[k + 1] - dz[k + 1];
        if (ix > 0) bt += bx[k] - dx[k];
        if (ix + 1 < N) bt -= bx[k + N] - dx[k + N];
        return g[k] - bt;
    };
    for (int e = threadIdx.x; e < C * N; e += blockDim.x) {
        const int c = e / N, i = e - c * N;
        const int ix = col0 + 2 * c;
        a0[c * N + nm_perm(i, N)] = make_double2(rhs(i, ix), rhs(i, ix + 1));
    }
    __syncthreads();
    const double2 *Z = nm_fft(a0, a1, N, C, A.f, A.tw, false);
    for (int e = threadIdx.x; e < C * N; e += blockDim.x) {
        const int c = e / N, k = e - c * N;
        const int ix = col0 + 2 * c;
        const double2 X = nm_dct_pair(Z + c * N, k, N, A.qw);
        if (ix < N) t[o + (size_t)ix * N + k] = X.x;
        if (ix + 1 < N) t[o + (size_t)(ix + 1) * N + k] = X.y;
    }
}

// Pass 2, along x: rows = pairs of z frequencies (kz, kz + 1); DCT-II along x, divide by
// 1 + l_kz + l_kx, DCT-III along x, in place in t.
__global__ __launch_bounds__(256) void k_nm_x(double *__restrict__ t, const int *__restrict__ active, NmArgs A)
{
    extern __shared__ double2 nm_sm[];
    const int b = blockIdx.y;
    if (!active[b]) return;
    const int N = A.N, C = A.C;
    const size_t o = (size_t)b * N * N;
    double2 *a0 = nm_sm, *a1 = nm_sm + C * N;
    const int row0 = blockIdx.x * 2 * C;
    const int R = 2 * C;
    // element (kz = row0 + r, ix): consecutive threads take consecutive kz
    for (int e = threadIdx.x; e < R * N; e += blockDim.x) {
        const int ix = e / R, r = e - ix * R;
        const int kz = row0 + r;
        const double v = kz < N ? t[o + (size_t)ix * N + kz] : 0.0;
        reinterpret_cast<double *>(a0 + (r >> 1) * N + nm_perm(ix, N))[r & 1] = v;
    }
    __syncthreads();
    double2 *Z = nm_fft(a0, a1, N, C, A.f, A.tw, false);
    double2 *X = Z == a0 ? a1 : a0;
    for (int e = threadIdx.x; e < C * N; e += blockDim.x) {
        const int c = e / N, k = e - c * N;
        const int kz = row0 + 2 * c;
        double2 v = nm_dct_pair(Z + c * N, k, N, A.qw);
        const double lk = A.lam[k];
        v.x /= 1.0 + A.lam[kz < N ? kz : 0] + lk;
        v.y /= 1.0 + A.lam[kz + 1 < N ? kz + 1 : 0] + lk;
        X[c * N + k] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < C * N; e += blockDim.x) {
        const int c = e / N, k = e - c * N;
        const double2 *Xc = X + c * N;
        Z[c * N + k] = nm_idct_pair(Xc[k], Xc[k == 0 ? 0 : N - k], k, N, A.qw);
    }
    __syncthreads();
    const double2 *z = nm_fft(Z, X, N, C, A.f, A.tw, true);
    const double inv = 1.0 / N;
    for (int e = threadIdx.x; e < R * N; e += blockDim.x) {
        const int ix = e / R, r = e - ix * R;
        const int kz = row0 + r;
        if (kz < N) t[o + (size_t)ix * N + kz] = reinterpret_cast<const double *>(z + (r >> 1) * N + nm_perm(ix, N))[r & 1] * inv;
    }
}

// Pass 3, along z: rows = pairs of image columns; DCT-III along z from t into u.
__global__ __launch_bounds__(256) void k_nm_z_inv(const double *__restrict__ t, double *__restrict__ u,
                                                  const int *__restrict__ active, NmArgs A)
{
    extern __shared__ double2 nm_sm[];
    const int b = blockIdx.y;
    if (!active[b]) return;
    const int N = A.N, C = A.C;
    const size_t o = (size_t)b * N * N;
    double2 *a0 = nm_sm, *a1 = nm_sm + C * N;
    const int col0 = blockIdx.x * 2 * C;
    for (int e = threadIdx.x; e < C * N; e += blockDim.x) {
        const int c = e / N, k = e - c * N;
        const int ix = col0 + 2 * c;
        const int kr = k == 0 ? 0 : N - k;
        double2 xk = make_double2(0.0, 0.0), xr = make_double2(0.0, 0.0);
        if (ix < N) { xk.x = t[o + (size_t)ix * N + k]; xr.x = t[o + (size_t)ix * N + kr]; }
        if (ix + 1 < N) { xk.y = t[o + (size_t)(ix + 1) * N + k]; xr.y = t[o + (size_t)(ix + 1) * N + kr]; }
        a0[c * N + k] = nm_idct_pair(xk, xr, k, N, A.qw);
    }
    __syncthreads();
    const double2 *z = nm_fft(a0, a1, N, C, A.f, A.tw, true);
    const double inv = 1.0 / N;
    for (int e = threadIdx.x; e < C * N; e += blockDim.x) {
        const int c = e / N, i = e - c * N;
        const int ix = col0 + 2 * c;
        const double2 v = z[c * N + nm_perm(i, N)];
        if (ix < N) u[o + (size_t)ix * N + i] = v.x * inv;
        if (ix + 1 < N) u[o + (size_t)(ix + 1) * N + i] = v.y * inv;
    }
}

// SB_ITV2's stop test on the device: an active item stops once
// ||u_prev - u|| / ||u|| <= 1e-3 (NaN also stops, as MATLAB's while condition); cnt counts
// the steps in which some item was active
__global__ void k_tv_active(const double *__restrict__ part, int *__restrict__ active, int B, int *cnt)
{
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        if (!active[b]) continue;
        any = 1;
        double sa = 0.0, sb = 0.0;
        for (int i = 0; i < kRedBlocks; ++i) {
            sa += part[((size_t)Q_A * B + b) * kRedBlocks + i];
            sb += part[((size_t)Q_B * B + b) * kRedBlocks + i];
        }
        const double err = sqrt(sa) / sqrt(sb);
        if (!(err > 1e-3)) active[b] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0 && any) cnt[0] += 1;
}

// Bub = B u + b; s = |Bub|; d = max(s - mu, 0) Bub / s (0 where s == 0); b = Bub - d;
// ||u_prev - u||^2 -> part Q_A, ||u||^2 -> Q_B; u_prev <- u
__global__ __launch_bounds__(256) void k_tv_shrink(const double *__restrict__ u, double *__restrict__ up,
                                                   double *__restrict__ bz, double *__restrict__ bx,
                                                   double *__restrict__ dz, double *__restrict__ dx, int N, int B,
                                                   const double *__restrict__ mu, const int *__restrict__ active,
                                                   double *part)
{
    const int b = blockIdx.y;
    const size_t P = (size_t)N * N, o = (size_t)b * P;
    double e = 0.0, nu = 0.0;
    if (active[b]) {
        const double m = mu[b];
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < P; i += (size_t)gridDim.x * blockDim.x) {
            const int ix = (int)(i / N), iz = (int)(i - (size_t)ix * N);
            const double ui = u[o + i];
            const double gz = iz > 0 ? ui - u[o + i - 1] : 0.0;
            const double gx = ix > 0 ? ui - u[o + i - N] : 0.0;
            const double az = gz + bz[o + i], ax = gx + bx[o + i];
            const double s = sqrt(az * az + ax * ax);
            const double f = s > 0.0 ? fmax(s - m, 0.0) / s : 0.0;
            const double nz = f * az, nx = f * ax;
            dz[o + i] = nz;
            dx[o + i] = nx;
            bz[o + i] = az - nz;
            bx[o + i] = ax - nx;
            const double df = up[o + i] - ui;
            e += df * df;
            nu += ui * ui;
            up[o + i] = ui;
        }
    }
    block_sum_store(e, part, Q_A, b, B);
    block_sum_store(nu, part, Q_B, b, B);
}

__global__ void k_tv_zero_cnt(int *cnt) { cnt[0] = 0; }

__global__ void k_tv_store(const double *__restrict__ u, float *__restrict__ x, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = (float)u[i];
}

// isotropic TV with MATLAB gradient (central differences inside, one-sided at edges) -> Q_C
__global__ __launch_bounds__(256) void k_itv(const float *__restrict__ x, int N, int B, double *part)
{
    const int b = blockIdx.y;
    const size_t P = (size_t)N * N, o = (size_t)b * P;
    double t = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < P; i += (size_t)gridDim.x * blockDim.x) {
        const int ix = (int)(i / N), iz = (int)(i - (size_t)ix * N);
        auto at = [&](int z_, int x_) { return (double)x[o + (size_t)x_ * N + z_]; };
        double gz, gx;
        if (N == 1) { gz = gx = 0.0; }
        else {
            gz = iz == 0 ? at(1, ix) - at(0, ix) : (iz == N - 1 ? at(N - 1, ix) - at(N - 2, ix)
                                                                 : 0.5 * (at(iz + 1, ix) - at(iz - 1, ix)));
            gx = ix == 0 ? at(iz, 1) - at(iz, 0) : (ix == N - 1 ? at(iz, N - 1) - at(iz, N - 2)
                                                                 : 0.5 * (at(iz, ix + 1) - at(iz, ix - 1)));
        }
        t += sqrt(gz * gz + gx * gx);
    }
    block_sum_store(t, part, Q_C, b, B);
}

// ---- SpaRSA vector kernels --------------------------------------------------------
struct TvConsts {
    float sl, sb;
    int coupled;
};

// resid_s = sl*(A w - y); resid_c = -sb*v; x = w; ||resid||^2 -> Q_D (sino) and Q_E (pix)
__global__ __launch_bounds__(256) void k_tv_init_s(const float *__restrict__ aw, const float *__restrict__ y,
                                                   float *__restrict__ rs, size_t n, int B, TvConsts k, double *part)
{
    const int b = blockIdx.y;
    const size_t o = (size_t)b * n;
    double e = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float r = k.sl * (aw[o + i] - y[o + i]);
        rs[o + i] = r;
        e += (double)r * r;
    }
    block_sum_store(e, part, Q_D, b, B);
}

__global__ __launch_bounds__(256) void k_tv_init_p(const float *__restrict__ w, const float *__restrict__ v,
                                                   float *__restrict__ rc, float *__restrict__ x, size_t n, int B,
                                                   TvConsts k, double *part)
{
    const int b = blockIdx.y;
    const size_t o = (size_t)b * n;
    double e = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float r = -k.sb * v[o + i];     // sb*w - sb*(w + v)
        rc[o + i] = r;
        x[o + i] = w[o + i];
        e += (double)r * r;
    }
    block_sum_store(e, part, Q_E, b, B);
}

// z = x - (sl*A'rs + sb*rc) / alpha_g ; prev_x = x
__global__ __launch_bounds__(256) void k_tv_step(const float *__restrict__ atr, const float *__restrict__ rc,
                                                 const float *__restrict__ x, float *__restrict__ z,
                                                 float *__restrict__ px, size_t n, int B, TvConsts k,
                                                 const double *__restrict__ inv_alpha, const int *__restrict__ gact)
{
    const int b = blockIdx.y;
    const int g = k.coupled ? 0 : b;
    if (!gact[g]) return;
    const size_t o = (size_t)b * n;
    const float ia = (float)inv_alpha[g];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float xi = x[o + i];
        px[o + i] = xi;
        z[o + i] = xi - (k.sl * atr[o + i] + k.sb * rc[o + i]) * ia;
    }
}

// dx = x - prev_x (x holds the prox output); ||dx||^2 -> Q_A; resid_c += sb*dx, ||resid_c||^2 -> Q_E
__global__ __launch_bounds__(256) void k_tv_dx(const float *__restrict__ x, const float *__restrict__ px,
                                               float *__restrict__ dxv, float *__restrict__ rc, size_t n, int B,
                                               TvConsts k, const int *__restrict__ gact, double *part)
{
    const int b = blockIdx.y;
    const int g = k.coupled ? 0 : b;
    const size_t o = (size_t)b * n;
    double dd = 0.0, ee = 0.0, aa = 0.0;
    if (gact[g]) {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const float d = x[o + i] - px[o + i];
            dxv[o + i] = d;
            dd += (double)d * d;
            const float adx = k.sb * d;
            aa += (double)adx * adx;
            const float r = rc[o + i] + adx;
            rc[o + i] = r;
            ee += (double)r * r;
        }
    }
    block_sum_store(dd, part, Q_A, b, B);
    block_sum_store(ee, part, Q_E, b, B);
    block_sum_store(aa, part, Q_F, b, B);
}

// resid_s += sl*(A dx); ||resid_s||^2 -> Q_D, ||sl A dx||^2 -> Q_B
__global__ __launch_bounds__(256) void k_tv_rs(const float *__restrict__ adx, float *__restrict__ rs, size_t n, int B,
                                               TvConsts k, const int *__restrict__ gact, double *part)
{
    const int b = blockIdx.y;
    const int g = k.coupled ? 0 : b;
    const size_t o = (size_t)b * n;
    double ee = 0.0, aa = 0.0;
    if (gact[g]) {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const float a = k.sl * adx[o + i];
            aa += (double)a * a;
            const float r = rs[o + i] + a;
            rs[o + i] = r;
            ee += (double)r * r;
        }
    }
    block_sum_store(ee, part, Q_D, b, B);
    block_sum_store(aa, part, Q_B, b, B);
}

__global__ void k_tv_sums(const double *part, int B, double *out)
{
    // out[q * B + b] = sum over blocks, q = 0..5
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * B) return;
    const int q = t / B, b = t - q * B;
    double s = 0.0;
    for (int i = 0; i < kRedBlocks; ++i) s += part[((size_t)q * B + b) * kRedBlocks + i];
    out[t] = s;
}

__global__ void k_tv_nnls(const float *__restrict__ x, float *__restrict__ w, float *__restrict__ v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float u = x[i], vi = v[i];
        const float wn = fmaxf(u - vi, 0.0f);
        w[i] = wn;
        v[i] = vi + wn - u;
    }
}

struct TvWork {
    std::vector<void *> all;
    ~TvWork() { for (void *p : all) (void)hipFree(p); }
    template <typename T>
    int alloc(T **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(MBREC_ERR_HIP, "mbrec:alloc", "TV workspace allocation failed");
        }
        all.push_back(*p);
        return MBREC_OK;
    }
};

static inline int nblk(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 4096); }

struct TvProx {
    int N = 0, B = 0;
    double2 *tw = nullptr, *qw = nullptr;          // e^{-2 pi i j/N}, e^{-i pi k/2N}
    double *lam = nullptr;                         // 4 sin^2(pi k / 2N)
    NmArgs nm{};
    double *g = nullptr, *u = nullptr, *up = nullptr, *bz = nullptr, *bx = nullptr, *dz = nullptr, *dx = nullptr;
    double *t1 = nullptr, *mu = nullptr, *part = nullptr, *sums = nullptr;
    int *active = nullptr, *cnt = nullptr;
};

// N's prime factors, radix 4 first (fewer stages), then 2, 3, 5, 7, ...
static bool nm_factors(int N, NmFactors &f)
{
    f.n = 0;
    int n = N;
    auto take = [&](int p) {
        while (n % p == 0) {
            if (f.n == kNmMaxFactors) return false;
            f.p[f.n++] = p;
            n /= p;
        }
        return true;
    };
    if (!take(4) || !take(2)) return false;
    for (int p = 3; n > 1; p += 2)
        if (!take(p)) return false;
    return true;
}

static int prox_setup(TvProx &T, TvWork &w, int N, int B, hipStream_t s)
{
    if (N > kNmMaxN)
        return set_error(MBREC_ERR_ARG, "mbrec:tvSize", "rec_nn_with_TV_reg: image side above 2048 is not supported");
    T.N = N;
    T.B = B;
    const size_t P = (size_t)N * N, n = P * B;
    int rc;
    double **bufs[] = {&T.g, &T.u, &T.up, &T.bz, &T.bx, &T.dz, &T.dx, &T.t1};
    for (double **b : bufs)
        if ((rc = w.alloc(b, n))) return rc;
    if ((rc = w.alloc(&T.tw, N)) || (rc = w.alloc(&T.qw, N)) || (rc = w.alloc(&T.lam, N)) ||
        (rc = w.alloc(&T.mu, B)) || (rc = w.alloc(&T.part, (size_t)kMaxQ * B * kRedBlocks)) ||
        (rc = w.alloc(&T.sums, 6 * (size_t)B)) || (rc = w.alloc(&T.active, B)) || (rc = w.alloc(&T.cnt, 1)))
        return rc;
    std::vector<double2> tw(N), qw(N);
    std::vector<double> lam(N);
    for (int k = 0; k < N; ++k) {
        lam[k] = 4.0 * std::pow(std::sin(M_PI * k / (2.0 * N)), 2);
        tw[k] = make_double2(std::cos(2.0 * M_PI * k / N), -std::sin(2.0 * M_PI * k / N));
        qw[k] = make_double2(std::cos(M_PI * k / (2.0 * N)), -std::sin(M_PI * k / (2.0 * N)));
    }
    MBREC_HIP(hipMemcpyAsync(T.tw, tw.data(), sizeof(double2) * N, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(T.qw, qw.data(), sizeof(double2) * N, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(T.lam, lam.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipStreamSynchronize(s));
    T.nm.N = N;
    // row pairs per workgroup: about 512 complex points (at 600^2 one pair, 19 KB of LDS,
    // eight workgroups per CU; measured 0.58 ms per Neumann solve of 32 images against
    // 1.00 ms with three pairs, 57.6 KB and two workgroups per CU)
    T.nm.C = std::max(1, std::min(4, 512 / N));
    if (!nm_factors(N, T.nm.f))
        return set_error(MBREC_ERR_ARG, "mbrec:tvSize", "rec_nn_with_TV_reg: too many prime factors in the image side");
    T.nm.tw = T.tw;
    T.nm.qw = T.qw;
    T.nm.lam = T.lam;
    return MBREC_OK;
}

// u = (I + B'B)^-1 (g - B'(b - d)) for every active item: DCT-II along z (rhs formed on
// load), along x (forward, divide by 1 + l_kz + l_kx, inverse), DCT-III along z
static void neumann_solve(TvProx &T, hipStream_t s)
{
    const int N = T.N, C = T.nm.C;
    const dim3 grid((unsigned)(((N + 1) / 2 + C - 1) / C), (unsigned)T.B);
    const size_t lds = sizeof(double2) * 2 * C * N;
    hipLaunchKernelGGL(k_nm_z_fwd, grid, dim3(256), lds, s, (const double *)T.g, (const double *)T.bz,
                       (const double *)T.bx, (const double *)T.dz, (const double *)T.dx, T.t1,
                       (const int *)T.active, T.nm);
    hipLaunchKernelGGL(k_nm_x, grid, dim3(256), lds, s, T.t1, (const int *)T.active, T.nm);
    hipLaunchKernelGGL(k_nm_z_inv, grid, dim3(256), lds, s, (const double *)T.t1, T.u, (const int *)T.active, T.nm);
}

// x (FP32, [B][N][N]) <- SB_ITV2(z, mu_b) per item; mu host [B]; items with skip[b] are left as they are.
// The stop test runs on the device after every split-Bregman step; the host reads the
// active flags every kTvCheck steps (an item that stopped inside a group is not updated
// again, so the result is the per-step loop's).
constexpr int kTvCheck = 4;

static int tv_prox(TvProx &T, const float *z, float *x, const std::vector<double> &mu,
                   const std::vector<int> &skip, hipStream_t s, int *iters_out)
{
    const int N = T.N, B = T.B;
    const size_t P = (size_t)N * N, n = P * B;
    std::vector<int> act(B);
    for (int b = 0; b < B; ++b) act[b] = skip[b] ? 0 : 1;
    MBREC_HIP(hipMemcpyAsync(T.mu, mu.data(), sizeof(double) * B, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(T.active, act.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_tv_load, dim3(nblk(n)), dim3(256), 0, s, z, T.g, T.up, T.bz, T.bx, T.dz, T.dx, n);
    hipLaunchKernelGGL(k_tv_zero_cnt, dim3(1), dim3(1), 0, s, T.cnt);
    for (;;) {
        bool any = false;
        for (int b = 0; b < B; ++b) any = any || act[b];
        if (!any) break;
        for (int c = 0; c < kTvCheck; ++c) {
            neumann_solve(T, s);
            hipLaunchKernelGGL(k_tv_shrink, vgrid(B), dim3(256), 0, s, (const double *)T.u, T.up, T.bz, T.bx, T.dz,
                               T.dx, N, B, (const double *)T.mu, (const int *)T.active, T.part);
            hipLaunchKernelGGL(k_tv_active, dim3(1), dim3(256), 0, s, (const double *)T.part, T.active, B, T.cnt);
        }
        MBREC_HIP(hipMemcpyAsync(act.data(), T.active, sizeof(int) * B, hipMemcpyDeviceToHost, s));
        MBREC_HIP(hipStreamSynchronize(s));
    }
    // write back items that were processed
    for (int b = 0; b < B; ++b)
        if (!skip[b])
            hipLaunchKernelGGL(k_tv_store, dim3(nblk(P)), dim3(256), 0, s, (const double *)(T.up + b * P), x + b * P, P);
    if (iters_out) {
        MBREC_HIP(hipMemcpyAsync(iters_out, T.cnt, sizeof(int), hipMemcpyDeviceToHost, s));
        MBREC_HIP(hipStreamSynchronize(s));
    }
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

int nnls_tv(Plan *p, const float *y, int B, int coupled, int num_iter, double tau, int inner, int inner_min,
            float *rec, double *lcurve, hipStream_t s)
{
    if (p->Nz != p->Nx)
        return set_error(MBREC_ERR_ARG, "rec_nn_with_TV_reg:square", "This TV implementation works only with square images!");
    const int N = p->Nz;
    const size_t np = (size_t)p->P, ns = (size_t)p->D * p->nT;
    const int G = coupled ? 1 : B;
    TvWork w;
    TvProx T;
    int rc;
    float *x, *px, *z, *dxv, *v, *rc_, *atr, *rs, *ts;
    double *part, *sums;
    float **pv[] = {&x, &px, &z, &dxv, &v, &rc_, &atr};
    for (float **q : pv)
        if ((rc = w.alloc(q, np * B))) return rc;
    if ((rc = w.alloc(&rs, ns * B)) || (rc = w.alloc(&ts, ns * B)) ||
        (rc = w.alloc(&part, (size_t)kMaxQ * B * kRedBlocks)) || (rc = w.alloc(&sums, 6 * (size_t)B)))
        return rc;
    if ((rc = prox_setup(T, w, N, B, s))) return rc;
    int *gact;
    double *inv_alpha;
    if ((rc = w.alloc(&gact, G)) || (rc = w.alloc(&inv_alpha, G))) return rc;
    TvConsts k{1.0f, (float)std::sqrt(1e-3), coupled};
    float *wv = rec;
    MBREC_HIP(hipMemsetAsync(wv, 0, sizeof(float) * np * B, s));
    MBREC_HIP(hipMemsetAsync(v, 0, sizeof(float) * np * B, s));
    std::vector<double> hs(6 * (size_t)B), f(G), prev_f(G), alpha(G), itv(B);
    std::vector<int> gactive(G), skip(B);
    auto group = [&](int b) { return coupled ? 0 : b; };
    auto reduce = [&]() -> int {
        hipLaunchKernelGGL(k_tv_sums, dim3((6 * B + 255) / 256), dim3(256), 0, s, (const double *)part, B, sums);
        MBREC_HIP(hipMemcpyAsync(hs.data(), sums, sizeof(double) * 6 * B, hipMemcpyDeviceToHost, s));
        MBREC_HIP(hipStreamSynchronize(s));
        return MBREC_OK;
    };
    auto phi = [&](const float *img) -> int {
        hipLaunchKernelGGL(k_itv, vgrid(B), dim3(256), 0, s, img, N, B, part);
        return reduce();
    };
    for (int it = 0; it < num_iter; ++it) {
        // SpaRSA init: x = w, resid = A_aug w - y_aug
        if ((rc = op_forward(p, wv, ts, B, true, s))) return rc;
        hipLaunchKernelGGL(k_tv_init_s, vgrid(B), dim3(256), 0, s, (const float *)ts, y, rs, ns, B, k, part);
        hipLaunchKernelGGL(k_tv_init_p, vgrid(B), dim3(256), 0, s, (const float *)wv, (const float *)v, rc_, x, np, B,
                           k, part);
        if ((rc = reduce())) return rc;
        std::vector<double> rr(G, 0.0);
        for (int b = 0; b < B; ++b) rr[group(b)] += hs[Q_D * B + b] + hs[Q_E * B + b];
        if ((rc = phi(x))) return rc;
        std::fill(f.begin(), f.end(), 0.0);
        for (int b = 0; b < B; ++b) f[group(b)] += tau * hs[Q_C * B + b];
        for (int g = 0; g < G; ++g) { f[g] += 0.5 * rr[g]; alpha[g] = 1.0; gactive[g] = 1; }
        // SpaRSA's iteration count: its counter is k + 1 after step k; it continues while
        // that is <= MiniterA, stops once it exceeds MaxiterA, and in between stops when the
        // relative change of the objective is <= ToleranceA (1e-8)
        const int steps = std::max(inner, inner_min);
        for (int kk = 0; kk < steps; ++kk) {
            // grad = A_aug' resid; z = x - grad / alpha
            if ((rc = op_transpose(p, rs, atr, B, true, s))) return rc;
            std::vector<double> ia(G);
            for (int g = 0; g < G; ++g) ia[g] = 1.0 / alpha[g];
            MBREC_HIP(hipMemcpyAsync(inv_alpha, ia.data(), sizeof(double) * G, hipMemcpyHostToDevice, s));
            MBREC_HIP(hipMemcpyAsync(gact, gactive.data(), sizeof(int) * G, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_tv_step, vgrid(B), dim3(256), 0, s, (const float *)atr, (const float *)rc_,
                               (const float *)x, z, px, np, B, k, (const double *)inv_alpha, (const int *)gact);
            // x = Psi(z, tau / alpha) per image
            std::vector<double> mu(B);
            for (int b = 0; b < B; ++b) {
                mu[b] = tau / alpha[group(b)];
                skip[b] = gactive[group(b)] ? 0 : 1;
            }
            if ((rc = tv_prox(T, z, x, mu, skip, s, nullptr))) return rc;
            // dx, A dx, resid update
            hipLaunchKernelGGL(k_tv_dx, vgrid(B), dim3(256), 0, s, (const float *)x, (const float *)px, dxv, rc_, np, B,
                               k, (const int *)gact, part);
            if ((rc = op_forward(p, dxv, ts, B, true, s))) return rc;
            hipLaunchKernelGGL(k_tv_rs, vgrid(B), dim3(256), 0, s, (const float *)ts, rs, ns, B, k, (const int *)gact,
                               part);
            if ((rc = reduce())) return rc;
            std::vector<double> dd(G, 0.0), adx(G, 0.0), res(G, 0.0);
            for (int b = 0; b < B; ++b) {
                const int g = group(b);
                dd[g] += hs[Q_A * B + b];
                adx[g] += hs[Q_B * B + b] + hs[Q_F * B + b];
                res[g] += hs[Q_D * B + b] + hs[Q_E * B + b];
            }
            if ((rc = phi(x))) return rc;
            std::vector<double> tv(G, 0.0);
            for (int b = 0; b < B; ++b) tv[group(b)] += hs[Q_C * B + b];
            for (int g = 0; g < G; ++g) {
                if (!gactive[g]) continue;
                prev_f[g] = f[g];
                f[g] = 0.5 * res[g] + tau * tv[g];
                double a = adx[g] / (2.2250738585072014e-308 + dd[g]);
                alpha[g] = std::min(1e30, std::max(1e-30, a));
                const int counter = kk + 2;
                if (counter <= inner_min) continue;
                if (counter > inner || std::fabs(f[g] - prev_f[g]) / prev_f[g] <= 1e-8) gactive[g] = 0;
            }
            bool any = false;
            for (int g = 0; g < G; ++g) any = any || gactive[g];
            if (!any) break;
        }
        hipLaunchKernelGGL(k_tv_nnls, dim3(nblk(np * B)), dim3(256), 0, s, (const float *)x, wv, v, np * B);
    }
    if (lcurve) {
        if ((rc = op_forward(p, wv, ts, B, true, s))) return rc;
        hipLaunchKernelGGL(k_tv_init_s, vgrid(B), dim3(256), 0, s, (const float *)ts, y, rs, ns, B, k, part);
        if ((rc = reduce())) return rc;
        std::vector<double> e(B);
        for (int b = 0; b < B; ++b) e[b] = hs[Q_D * B + b];
        if ((rc = phi(wv))) return rc;
        for (int g = 0; g < G; ++g) { lcurve[2 * g] = 0.0; lcurve[2 * g + 1] = 0.0; }
        for (int b = 0; b < B; ++b) {
            lcurve[2 * group(b)] += e[b];
            lcurve[2 * group(b) + 1] += hs[Q_C * B + b];
        }
    }
    MBREC_HIP(hipStreamSynchronize(s));
    return MBREC_OK;
}

}  // namespace mbrec

using namespace mbrec;

extern "C" int mbrec_rec_nn_with_TV_reg_dev(mbrec_plan h, const float *data, int B, int coupled, int num_iter,
                                            double regParam, int inner_iterations, int inner_min_iterations,
                                            float *rec, double *lcurve, mbrec_stream s)
{
    if (!h || !data || !rec || B <= 0 || num_iter < 0 || inner_iterations <= 0 || inner_min_iterations < 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "rec_nn_with_TV_reg: invalid arguments");
    return nnls_tv(&h->p, data, B, coupled, num_iter, regParam, inner_iterations, inner_min_iterations, rec, lcurve,
                   (hipStream_t)s);
}

extern "C" int mbrec_sb_itv2_dev(const float *g, int N, int B, const double *mu, float *u, int *iterations,
                                 mbrec_stream s)
{
    if (!g || !u || !mu || N <= 0 || B <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "SB_ITV2: invalid arguments");
    TvWork w;
    TvProx T;
    const hipStream_t st = (hipStream_t)s;
    int rc;
    if ((rc = prox_setup(T, w, N, B, st))) return rc;
    std::vector<double> m(mu, mu + B);
    std::vector<int> skip(B, 0);
    if ((rc = tv_prox(T, g, u, m, skip, st, iterations))) return rc;
    MBREC_HIP(hipStreamSynchronize(st));
    return MBREC_OK;
}

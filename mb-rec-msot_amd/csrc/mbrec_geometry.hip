// mbrec_geometry.hip -- travel-time tables on the device (MSOTmodel.m:350-382).
//
// For every (pixel p, detector q): refraction.m:20-24 -- a pixel on the detector side of
// the membrane (z_p <= m(x_p), always so for a single speed of sound, whose membrane is
// at +inf) travels straight, t = |q - p| / c_above; otherwise fermat_principle.m:12-32 --
// t = min over x in [-R, R] of |p - (x, m(x))| / c_below + |(x, m(x)) - q| / c_above, and
// the minimiser x is the membrane intersection.  The reference minimises with fminbnd
// (Brent's method, TolX 1e-4 m); here the minimum is found to about 1e-12 m: a scan of
// kScan points brackets the smallest sample, then golden-section search narrows the
// bracket.  The membrane is a curve z = m(x) given by samples on a uniform x grid
// (linear interpolation, clamped at the ends).  Travel-time indices follow
// MSOTmodel.m:380-381: tau = t fs - delayBeforeRecording - numCroppedSamplesAtSinogramStart.
#include "mbrec_internal.h"

#include <cmath>
#include <cstring>

namespace mbrec {

constexpr int kScan = 128;

struct Membrane {
    const double *z = nullptr;   // samples m(x0 + i dx), i < n; null: no membrane (single speed of sound)
    int n = 0;
    double x0 = 0.0, dx = 1.0;
};

__device__ __forceinline__ double membrane_at(const Membrane &m, double x)
{
    double u = (x - m.x0) / m.dx;
    if (u <= 0.0) return m.z[0];
    if (u >= m.n - 1) return m.z[m.n - 1];
    const int i = (int)floor(u);
    const double f = u - i;
    return m.z[i] + f * (m.z[i + 1] - m.z[i]);
}

__device__ __forceinline__ double path_time(const Membrane &m, double x, double xp, double zp, double xq, double zq,
                                            double ib, double ia)
{
    const double zm = membrane_at(m, x);
    const double a = xp - x, b = zp - zm, c = x - xq, d = zm - zq;
    return sqrt(a * a + b * b) * ib + sqrt(c * c + d * d) * ia;
}

__global__ __launch_bounds__(256) void k_travel_times(const double *__restrict__ xp_, const double *__restrict__ zp_,
                                                      int P, const double *__restrict__ xq_,
                                                      const double *__restrict__ zq_, int D, double c_below,
                                                      double c_above, double radius, Membrane m, double fs,
                                                      double shift, double *__restrict__ tau,
                                                      double *__restrict__ xint)
{
    const size_t n = (size_t)P * D;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int d = (int)(i / P), x = (int)(i - (size_t)d * P);
        const double xp = xp_[x], zp = zp_[x], xq = xq_[d], zq = zq_[d];
        double t, xi = __builtin_nan("");
        if (!m.z || !(zp > membrane_at(m, xp))) {
            const double a = xq - xp, b = zq - zp;
            t = sqrt(a * a + b * b) / c_above;
        } else {
            const double ib = 1.0 / c_below, ia = 1.0 / c_above;
            const double h = 2.0 * radius / (kScan - 1);
            int best = 0;
            double fb = path_time(m, -radius, xp, zp, xq, zq, ib, ia);
            for (int k = 1; k < kScan; ++k) {
                const double f = path_time(m, -radius + k * h, xp, zp, xq, zq, ib, ia);
                if (f < fb) { fb = f; best = k; }
            }
            double lo = -radius + fmax(best - 1, 0) * h, hi = -radius + fmin(best + 1, kScan - 1) * h;
            const double g = 0.6180339887498949;
            double x1 = hi - g * (hi - lo), x2 = lo + g * (hi - lo);
            double f1 = path_time(m, x1, xp, zp, xq, zq, ib, ia), f2 = path_time(m, x2, xp, zp, xq, zq, ib, ia);
            for (int it = 0; it < 64 && hi - lo > 1e-13; ++it) {
                if (f1 <= f2) {
                    hi = x2; x2 = x1; f2 = f1;
                    x1 = hi - g * (hi - lo);
                    f1 = path_time(m, x1, xp, zp, xq, zq, ib, ia);
                } else {
                    lo = x1; x1 = x2; f1 = f2;
                    x2 = lo + g * (hi - lo);
                    f2 = path_time(m, x2, xp, zp, xq, zq, ib, ia);
                }
            }
            xi = 0.5 * (lo + hi);
            t = fmin(fb, path_time(m, xi, xp, zp, xq, zq, ib, ia));
        }
        tau[i] = t * fs - shift;
        if (xint) xint[i] = xi;
    }
}

}  // namespace mbrec

using namespace mbrec;

extern "C" int mbrec_travel_times_dev(const double *xpix, const double *zpix, int P, const double *xdet,
                                      const double *zdet, int D, double c_below, double c_above, double radius,
                                      const double *membrane_z, int nm, double membrane_x0, double membrane_dx,
                                      double fs, double delay, double ncrop, double *tau, double *x_intersections,
                                      mbrec_stream stream)
{
    if (!xpix || !zpix || !xdet || !zdet || !tau || P <= 0 || D <= 0 || !(c_above > 0) || !(fs > 0) ||
        (membrane_z && (nm < 2 || !(membrane_dx > 0) || !(c_below > 0) || !(radius > 0))))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "travel_times: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    double *dq = nullptr, *dm = nullptr;
    MBREC_HIP(hipMallocAsync((void **)&dq, sizeof(double) * 2 * D, s));
    MBREC_HIP(hipMemcpyAsync(dq, xdet, sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(dq + D, zdet, sizeof(double) * D, hipMemcpyHostToDevice, s));
    Membrane m;
    if (membrane_z) {
        MBREC_HIP(hipMallocAsync((void **)&dm, sizeof(double) * nm, s));
        MBREC_HIP(hipMemcpyAsync(dm, membrane_z, sizeof(double) * nm, hipMemcpyHostToDevice, s));
        m.z = dm;
        m.n = nm;
        m.x0 = membrane_x0;
        m.dx = membrane_dx;
    }
    const size_t n = (size_t)P * D;
    hipLaunchKernelGGL(k_travel_times, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)), dim3(256), 0, s, xpix,
                       zpix, P, (const double *)dq, (const double *)(dq + D), D, c_below, c_above, radius, m, fs,
                       delay + ncrop, tau, x_intersections);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipFreeAsync(dq, s));
    if (dm) MBREC_HIP(hipFreeAsync(dm, s));
    MBREC_HIP(hipStreamSynchronize(s));   // xdet / zdet / membrane_z are host buffers
    return MBREC_OK;
}

// ---------------------------------------------------------------------------
// SIR (SPMR) tables in the reference's polar layout (MSOTmodel.m:463-616).
//
// mbrec_spmr_grid_dev: per (pixel, detector) the distance and azimuth angle
// (MSOTmodel.m:594-596: acos(((q - p) . q) / (|q| |q - p|)), the cosine clipped to [-1, 1])
// and, after the grid is fixed (calculate_required_SPMR_distances_and_angels,
// MSOTmodel.m:599-616), the 1-based index (round(1 + (dist - dmin) / dd) - 1) n_angle +
// 1 + round(angle_deg / da) (MSOTmodel.m:472-478, MATLAB round).
//
// mbrec_spmr_table_dev: the table itself.  The reference computes it with Field II (a
// closed binary, not used); the generator here splits the element's in-plane width into
// n_sub point sub-elements, deposits each one's arrival-time offset relative to the
// element centre as a linear hat on the Ls-sample grid (centre Ls / 2), convolves with the
// pixel-model N-shape taps, and, as calculate_polarSPMR.m:152-161 does per entry, zeroes
// the samples below trim x the entry's largest magnitude and marks the first / last
// non-zero sample as its window (1-based, inclusive).
// ---------------------------------------------------------------------------
namespace mbrec {

__device__ __forceinline__ double matlab_round(double v) { return v >= 0 ? floor(v + 0.5) : -floor(-v + 0.5); }

__device__ __forceinline__ void polar_of(double xp, double zp, double xq, double zq, double &dist, double &ang)
{
    const double a = xp - xq, b = zp - zq;
    dist = sqrt(a * a + b * b);
    double cs = ((xq - xp) * xq + (zq - zp) * zq) / (sqrt(xq * xq + zq * zq) * dist);
    cs = fmin(1.0, fmax(-1.0, cs));
    ang = acos(cs);
}

// block-wide min / max of dist and max of angle -> atomics on the FP64 bit patterns
// (non-negative values order like their bits)
__global__ __launch_bounds__(256) void k_spmr_extent(const double *__restrict__ xv, const double *__restrict__ zv,
                                                     int per_det, int P, const double *__restrict__ xq,
                                                     const double *__restrict__ zq, int D,
                                                     unsigned long long *ext)
{
    __shared__ double smin[256], smax[256], samax[256];
    double dmin = 1e300, dmax = 0.0, amax = 0.0;
    const size_t n = (size_t)P * D;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int d = (int)(i / P), x = (int)(i - (size_t)d * P);
        const size_t at = per_det ? i : (size_t)x;
        double dist, ang;
        polar_of(xv[at], zv[at], xq[d], zq[d], dist, ang);
        dmin = fmin(dmin, dist);
        dmax = fmax(dmax, dist);
        amax = fmax(amax, ang);
    }
    smin[threadIdx.x] = dmin;
    smax[threadIdx.x] = dmax;
    samax[threadIdx.x] = amax;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + o]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
            samax[threadIdx.x] = fmax(samax[threadIdx.x], samax[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMin(&ext[0], (unsigned long long)__double_as_longlong(smin[0]));
        atomicMax(&ext[1], (unsigned long long)__double_as_longlong(smax[0]));
        atomicMax(&ext[2], (unsigned long long)__double_as_longlong(samax[0]));
    }
}

__global__ __launch_bounds__(256) void k_spmr_index(const double *__restrict__ xv, const double *__restrict__ zv,
                                                    int per_det, int P, const double *__restrict__ xq,
                                                    const double *__restrict__ zq, int D, double dmin, double dd,
                                                    double da, int n_angle, int32_t *__restrict__ sidx)
{
    const double r2d = 180.0 / 3.14159265358979323846;
    const size_t n = (size_t)P * D;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int d = (int)(i / P), x = (int)(i - (size_t)d * P);
        const size_t at = per_det ? i : (size_t)x;
        double dist, ang;
        polar_of(xv[at], zv[at], xq[d], zq[d], dist, ang);
        const double i_d = matlab_round(1.0 + (dist - dmin) / dd);
        const double i_a = 1.0 + matlab_round(ang * r2d / da);
        sidx[i] = (int32_t)((i_d - 1.0) * n_angle + i_a);
    }
}

constexpr int kMaxSub = 32;

// table entry e = (distance di, angle ai), angle fastest; out[e][t], t < Ls
__global__ __launch_bounds__(256) void k_spmr_table(const double *__restrict__ dists, const double *__restrict__ angs,
                                                    int nd, int na, double width, int n_sub, double c, double fs,
                                                    const double *__restrict__ nsh, int Ln, int off, int Ls,
                                                    double *__restrict__ out)
{
    const double d2r = 3.14159265358979323846 / 180.0;
    const size_t ne = (size_t)nd * na;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (size_t)gridDim.x * blockDim.x) {
        const double rho = dists[e / na], sa = sin(angs[e % na] * d2r);
        const int centre = Ls / 2;
        int lo[kMaxSub];
        double fr[kMaxSub];
        for (int s = 0; s < n_sub; ++s) {
            const double ei = (s + 0.5) * width / n_sub - width / 2.0;
            const double di = sqrt(rho * rho + ei * ei - 2.0 * rho * ei * sa);
            const double dt = (di - rho) / c * fs + centre;
            const double f = floor(dt);
            lo[s] = (f >= 0.0 && f + 1.0 < Ls) ? (int)f : -1;   // both hat samples inside, as deposited
            fr[s] = dt - f;
        }
        for (int t = 0; t < Ls; ++t) {
            double acc = 0.0;
            for (int s = 0; s < n_sub; ++s) {
                if (lo[s] < 0) continue;
                const int k0 = t - lo[s] - off, k1 = k0 - 1;
                if (k0 >= 0 && k0 < Ln) acc += nsh[k0] * ((1.0 - fr[s]) / n_sub);
                if (k1 >= 0 && k1 < Ln) acc += nsh[k1] * (fr[s] / n_sub);
            }
            out[e * Ls + t] = acc;
        }
    }
}

// calculate_polarSPMR.m:152-161 per entry: magnitudes below trim x the entry's largest set
// to zero, the window = first / last non-zero sample (1-based; an all-zero entry (1, 0))
__global__ __launch_bounds__(256) void k_spmr_windows(double *__restrict__ tab, size_t ne, int Ls, double trim,
                                                      int32_t *__restrict__ se)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (size_t)gridDim.x * blockDim.x) {
        double *v = tab + e * Ls;
        double mx = 0.0;
        for (int t = 0; t < Ls; ++t) mx = fmax(mx, fabs(v[t]));
        int first = -1, last = -1;
        for (int t = 0; t < Ls; ++t) {
            if (fabs(v[t]) < mx * trim) v[t] = 0.0;
            if (v[t] != 0.0) {
                if (first < 0) first = t;
                last = t;
            }
        }
        se[2 * e] = first < 0 ? 1 : first + 1;
        se[2 * e + 1] = first < 0 ? 0 : last + 1;
    }
}

}  // namespace mbrec

extern "C" int mbrec_spmr_grid_dev(const double *xv, const double *zv, int per_detector, int P, const double *xdet,
                                   const double *zdet, int D, double resolution_fov, double c, double *grid,
                                   int32_t *sidx, mbrec_stream stream)
{
    if (!xv || !zv || !xdet || !zdet || !grid || P <= 0 || D <= 0 || !(resolution_fov > 0) || !(c > 0))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "spmr_grid: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    double *dq = nullptr;
    unsigned long long *ext = nullptr;
    MBREC_HIP(hipMallocAsync((void **)&dq, sizeof(double) * 2 * D, s));
    MBREC_HIP(hipMallocAsync((void **)&ext, sizeof(unsigned long long) * 3, s));
    MBREC_HIP(hipMemcpyAsync(dq, xdet, sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(dq + D, zdet, sizeof(double) * D, hipMemcpyHostToDevice, s));
    const double big = 1e300;
    unsigned long long init[3];
    std::memcpy(&init[0], &big, 8);
    init[1] = init[2] = 0ull;
    MBREC_HIP(hipMemcpyAsync(ext, init, sizeof(init), hipMemcpyHostToDevice, s));
    const size_t n = (size_t)P * D;
    const dim3 g((unsigned)std::min<size_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_spmr_extent, g, dim3(256), 0, s, xv, zv, per_detector, P, (const double *)dq,
                       (const double *)(dq + D), D, ext);
    unsigned long long e[3];
    MBREC_HIP(hipMemcpyAsync(e, ext, sizeof(e), hipMemcpyDeviceToHost, s));
    MBREC_HIP(hipStreamSynchronize(s));
    double dmin_p, dmax_p, amax_p;
    std::memcpy(&dmin_p, &e[0], 8);
    std::memcpy(&dmax_p, &e[1], 8);
    std::memcpy(&amax_p, &e[2], 8);
    // calculate_required_SPMR_distances_and_angels (MSOTmodel.m:599-616)
    const double pi = 3.14159265358979323846;
    const double dd = resolution_fov * (1450.0 / c);
    const double dmin = dmin_p - dd, dmax = dmax_p + dd;
    const double da = std::atan2(dd, dmax) / pi * 180.0;
    const double amax = std::ceil((amax_p / pi * 180.0) * da) / da;
    // table axes as calculate_polarSPMR.m:22-25 lays them out
    const int n_dist = (int)std::ceil((dmax - dmin) / dd) + 1;
    const int n_angle = (int)std::ceil(amax / da) + 1;
    grid[0] = dmin; grid[1] = dmax; grid[2] = dd; grid[3] = amax; grid[4] = da;
    grid[5] = n_dist; grid[6] = n_angle;
    if (sidx)
        hipLaunchKernelGGL(k_spmr_index, g, dim3(256), 0, s, xv, zv, per_detector, P, (const double *)dq,
                           (const double *)(dq + D), D, dmin, dd, da, n_angle, sidx);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipFreeAsync(dq, s));
    MBREC_HIP(hipFreeAsync(ext, s));
    MBREC_HIP(hipStreamSynchronize(s));
    return MBREC_OK;
}

extern "C" int mbrec_spmr_index_dev(const double *xv, const double *zv, int per_detector, int P, const double *xdet,
                                    const double *zdet, int D, double dmin, double dd, double da, int n_angle,
                                    int32_t *sidx, mbrec_stream stream)
{
    if (!xv || !zv || !xdet || !zdet || !sidx || P <= 0 || D <= 0 || !(dd > 0) || !(da > 0) || n_angle <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "spmr_index: invalid arguments");
    hipStream_t s = (hipStream_t)stream;
    double *dq = nullptr;
    MBREC_HIP(hipMallocAsync((void **)&dq, sizeof(double) * 2 * D, s));
    MBREC_HIP(hipMemcpyAsync(dq, xdet, sizeof(double) * D, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(dq + D, zdet, sizeof(double) * D, hipMemcpyHostToDevice, s));
    const size_t n = (size_t)P * D;
    hipLaunchKernelGGL(k_spmr_index, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, s, xv, zv,
                       per_detector, P, (const double *)dq, (const double *)(dq + D), D, dmin, dd, da, n_angle, sidx);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipFreeAsync(dq, s));
    MBREC_HIP(hipStreamSynchronize(s));
    return MBREC_OK;
}

extern "C" int mbrec_spmr_table_dev(const double *dists, int n_dist, const double *angles_deg, int n_angle,
                                    double element_width, int n_sub, double c, double fs, const double *nshape,
                                    int Ln, int nshape_offset, int Ls, double trim, double *table, int32_t *windows,
                                    mbrec_stream stream)
{
    if (!dists || !angles_deg || !nshape || !table || !windows || n_dist <= 0 || n_angle <= 0 || n_sub <= 0 ||
        n_sub > kMaxSub || Ln <= 0 || Ls <= 1 || !(c > 0) || !(fs > 0))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "spmr_table: invalid arguments (at most %d sub-elements)",
                         kMaxSub);
    hipStream_t s = (hipStream_t)stream;
    double *buf = nullptr;
    MBREC_HIP(hipMallocAsync((void **)&buf, sizeof(double) * (n_dist + n_angle + Ln), s));
    MBREC_HIP(hipMemcpyAsync(buf, dists, sizeof(double) * n_dist, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(buf + n_dist, angles_deg, sizeof(double) * n_angle, hipMemcpyHostToDevice, s));
    MBREC_HIP(hipMemcpyAsync(buf + n_dist + n_angle, nshape, sizeof(double) * Ln, hipMemcpyHostToDevice, s));
    const size_t ne = (size_t)n_dist * n_angle;
    const dim3 g((unsigned)std::min<size_t>((ne + 255) / 256, 8192));
    hipLaunchKernelGGL(k_spmr_table, g, dim3(256), 0, s, (const double *)buf, (const double *)(buf + n_dist), n_dist,
                       n_angle, element_width, n_sub, c, fs, (const double *)(buf + n_dist + n_angle), Ln,
                       nshape_offset, Ls, table);
    hipLaunchKernelGGL(k_spmr_windows, g, dim3(256), 0, s, table, ne, Ls, trim, windows);
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipFreeAsync(buf, s));
    MBREC_HIP(hipStreamSynchronize(s));   // dists / angles / nshape are host buffers
    return MBREC_OK;
}

// mbrec_solver_l1.hip -- device-resident rec_nn_with_L1_eye_reg.
//
// nnls_L1_eye_matrix_less (reconstruction/nnls_L1_eye_matrix_less.m:24-55),
// split Bregman with x >= 0: per outer iteration
//     d = P(u - b);  b_in = P(d) + b + lam A'y - lam eta;  u = cgs(lam A'A + I, b_in, 1e-8, 2);
//     b = b + P(d) - u
// around CGS with MATLAB cgs semantics (Sonneveld 1989; x0 = 0, the iterate of
// minimal residual is tracked and returned when its recomputed residual is
// not larger than the last one).  A'y is computed once: the reference
// recomputes the identical value every outer iteration (:43).
//
// Same FP32-vector / FP64-scalar split and device-side scalar recurrences as
// the L2 solver.  One deviation from MATLAB's cgs is documented: when the
// recursive residual falls below tol*||b|| the iteration stops on that
// estimate, whereas MATLAB first recomputes b - A x; at tol = 1e-8 and 2
// iterations this branch is not reached in practice.
#include "mbrec_reduce.h"

#include <cmath>
#include <vector>

namespace mbrec {

struct CgsScal {
    double rho, rho1, beta, alpha, normr, normr_act, normrmin, tolb, n2b;
    int stag, flag, active, take_min;
};

struct L1Consts {
    float lam, lam_eta;
    double tol;
    int coupled;
};

static __device__ __forceinline__ float pos(float v) { return v > 0.0f ? v : 0.0f; }

// d = P(u - b); b_in = P(d) + b + lam*aty - lam*eta; r = rt = b_in; x = xmin = 0;  ||b_in||^2 -> Q_A
__global__ __launch_bounds__(256) void k_l1_setup(const float *__restrict__ u, const float *__restrict__ bp,
                                                  const float *__restrict__ aty, float *__restrict__ dp,
                                                  float *__restrict__ bin, float *__restrict__ r,
                                                  float *__restrict__ rt, float *__restrict__ x,
                                                  float *__restrict__ xmin, size_t n, int B, L1Consts k,
                                                  double *part)
{
    const int b = blockIdx.y;
    const size_t off = (size_t)b * n;
    double bb = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float d = pos(u[off + i] - bp[off + i]);
        dp[off + i] = d;
        const float v = pos(d) + bp[off + i] + k.lam * aty[off + i] - k.lam_eta;
        bin[off + i] = v;
        r[off + i] = v;
        rt[off + i] = v;
        x[off + i] = 0.0f;
        xmin[off + i] = 0.0f;
        bb += (double)v * v;
    }
    block_sum_store(bb, part, Q_A, b, B);
}

__global__ void k_cgs_init(CgsScal *sc, const double *part, int G, int B, L1Consts k)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        CgsScal &S = sc[g];
        const double bb = group_sum(part, Q_A, g, B, k.coupled);
        S.n2b = sqrt(bb);
        S.tolb = k.tol * S.n2b;
        S.normr = S.n2b;          // x0 = 0: r = b
        S.normr_act = S.normr;
        S.normrmin = S.normr;
        S.rho = 1.0;
        S.stag = 0;
        S.flag = 1;
        S.take_min = 0;
        S.active = 1;
        if (S.n2b == 0.0 || S.normr <= S.tolb) { S.flag = 0; S.active = 0; }
    }
}

// rho partial: rt'r -> Q_B
__global__ __launch_bounds__(256) void k_dot(const float *__restrict__ a, const float *__restrict__ c, size_t n,
                                             int B, int q, double *part)
{
    const int b = blockIdx.y;
    const size_t off = (size_t)b * n;
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += (double)a[off + i] * c[off + i];
    block_sum_store(s, part, q, b, B);
}

__global__ void k_cgs_rho(CgsScal *sc, const double *part, int G, int B, L1Consts k, int ii)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        CgsScal &S = sc[g];
        if (!S.active) continue;
        S.rho1 = S.rho;
        S.rho = group_sum(part, Q_B, g, B, k.coupled);
        if (S.rho == 0.0 || isinf(S.rho)) { S.flag = 4; S.active = 0; continue; }
        if (ii > 1) {
            S.beta = S.rho / S.rho1;
            if (S.beta == 0.0 || isinf(S.beta)) { S.flag = 4; S.active = 0; }
        }
    }
}

// ii == 1: uc = r, p = uc;  else uc = r + beta q, p = uc + beta (q + beta p)
__global__ __launch_bounds__(256) void k_cgs_dir(const float *__restrict__ r, const float *__restrict__ q,
                                                 float *__restrict__ uc, float *__restrict__ p, size_t n, int B,
                                                 L1Consts k, const CgsScal *sc, int ii)
{
    const int b = blockIdx.y;
    const CgsScal &S = sc[k.coupled ? 0 : b];
    if (!S.active) return;
    const size_t off = (size_t)b * n;
    const float be = (float)S.beta;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (ii == 1) {
            uc[off + i] = r[off + i];
            p[off + i] = r[off + i];
        } else {
            const float u = r[off + i] + be * q[off + i];
            uc[off + i] = u;
            p[off + i] = u + be * (q[off + i] + be * p[off + i]);
        }
    }
}

// out = lam*atav + v (the operator lam A'A + I); optional dot with rt -> q
__global__ __launch_bounds__(256) void k_l1_apply(const float *__restrict__ atav, const float *__restrict__ v,
                                                  float *__restrict__ out, const float *__restrict__ rt,
                                                  size_t n, int B, L1Consts k, int q, double *part)
{
    const int b = blockIdx.y;
    const size_t off = (size_t)b * n;
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float o = k.lam * atav[off + i] + v[off + i];
        out[off + i] = o;
        if (rt) s += (double)rt[off + i] * o;
    }
    if (rt) block_sum_store(s, part, q, b, B);
}

__global__ void k_cgs_alpha(CgsScal *sc, const double *part, int G, int B, L1Consts k)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        CgsScal &S = sc[g];
        if (!S.active) continue;
        const double rtvh = group_sum(part, Q_C, g, B, k.coupled);
        if (rtvh == 0.0) { S.flag = 4; S.active = 0; continue; }
        S.alpha = S.rho / rtvh;
        if (isinf(S.alpha)) { S.flag = 4; S.active = 0; }
    }
}

// q = uc - alpha vh; uh = uc + q; ||uh||^2 -> Q_D, ||x_old||^2 -> Q_E; x += alpha uh
__global__ __launch_bounds__(256) void k_cgs_q(const float *__restrict__ uc, const float *__restrict__ vh,
                                               float *__restrict__ q, float *__restrict__ uh,
                                               float *__restrict__ x, size_t n, int B, L1Consts k,
                                               const CgsScal *sc, double *part)
{
    const int b = blockIdx.y;
    const CgsScal &S = sc[k.coupled ? 0 : b];
    const size_t off = (size_t)b * n;
    double su = 0.0, sx = 0.0;
    if (S.active) {
        const float al = (float)S.alpha;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const float qq = uc[off + i] - al * vh[off + i];
            q[off + i] = qq;
            const float u = uc[off + i] + qq;
            uh[off + i] = u;
            su += (double)u * u;
            const float xo = x[off + i];
            sx += (double)xo * xo;
            x[off + i] = xo + al * u;
        }
    }
    block_sum_store(su, part, Q_D, b, B);
    block_sum_store(sx, part, Q_E, b, B);
}

// r -= alpha qh (qh = lam A'A uh + uh); ||r||^2 -> Q_F
__global__ __launch_bounds__(256) void k_cgs_r(const float *__restrict__ atauh, const float *__restrict__ uh,
                                               float *__restrict__ r, size_t n, int B, L1Consts k,
                                               const CgsScal *sc, double *part)
{
    const int b = blockIdx.y;
    const CgsScal &S = sc[k.coupled ? 0 : b];
    const size_t off = (size_t)b * n;
    double s = 0.0;
    if (S.active) {
        const float al = (float)S.alpha;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const float qh = k.lam * atauh[off + i] + uh[off + i];
            const float rn = r[off + i] - al * qh;
            r[off + i] = rn;
            s += (double)rn * rn;
        }
    }
    block_sum_store(s, part, Q_F, b, B);
}

__global__ void k_cgs_post(CgsScal *sc, const double *part, int G, int B, L1Consts k)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        CgsScal &S = sc[g];
        S.take_min = 0;
        if (!S.active) continue;
        const double uu = group_sum(part, Q_D, g, B, k.coupled);
        const double xx = group_sum(part, Q_E, g, B, k.coupled);
        if (fabs(S.alpha) * sqrt(uu) < 2.220446049250313e-16 * sqrt(xx)) S.stag++;
        else S.stag = 0;
        S.normr = sqrt(group_sum(part, Q_F, g, B, k.coupled));
        S.normr_act = S.normr;
        if (S.normr <= S.tolb) { S.flag = 0; S.active = 0; continue; }
        if (S.normr_act < S.normrmin) { S.normrmin = S.normr_act; S.take_min = 1; }
    }
}

// xmin = x where the scalar step recorded a new minimum
__global__ __launch_bounds__(256) void k_cgs_keepmin(const float *__restrict__ x, float *__restrict__ xmin,
                                                     size_t n, int B, L1Consts k, const CgsScal *sc)
{
    const int b = blockIdx.y;
    const CgsScal &S = sc[k.coupled ? 0 : b];
    if (!S.take_min) return;
    const size_t off = (size_t)b * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        xmin[off + i] = x[off + i];
}

// r_comp = b_in - (lam A'A xmin + xmin); ||.||^2 -> Q_A
__global__ __launch_bounds__(256) void k_cgs_rcomp(const float *__restrict__ bin, const float *__restrict__ ataxm,
                                                   const float *__restrict__ xmin, size_t n, int B, L1Consts k,
                                                   double *part)
{
    const int b = blockIdx.y;
    const size_t off = (size_t)b * n;
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float r = bin[off + i] - (k.lam * ataxm[off + i] + xmin[off + i]);
        s += (double)r * r;
    }
    block_sum_store(s, part, Q_A, b, B);
}

__global__ void k_cgs_select(CgsScal *sc, const double *part, int G, int B, L1Consts k)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        CgsScal &S = sc[g];
        S.take_min = 0;
        if (S.flag == 0) continue;
        const double rc = sqrt(group_sum(part, Q_A, g, B, k.coupled));
        if (rc <= S.normr_act) S.take_min = 1;
    }
}

// u = (take_min ? xmin : x); b += P(d) - u
__global__ __launch_bounds__(256) void k_l1_update(const float *__restrict__ x, const float *__restrict__ xmin,
                                                   const float *__restrict__ dp, float *__restrict__ u,
                                                   float *__restrict__ bp, size_t n, int B, L1Consts k,
                                                   const CgsScal *sc)
{
    const int b = blockIdx.y;
    const CgsScal &S = sc[k.coupled ? 0 : b];
    const size_t off = (size_t)b * n;
    const float *src = S.take_min ? xmin : x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float un = src[off + i];
        u[off + i] = un;
        bp[off + i] = bp[off + i] + pos(dp[off + i]) - un;
    }
}

__global__ __launch_bounds__(256) void k_zero(float *__restrict__ a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0.0f;
}

struct L1Work {
    std::vector<void *> all;
    ~L1Work()
    {
        for (void *p : all) (void)hipFree(p);
    }
    template <typename T>
    int alloc(T **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(MBREC_ERR_HIP, "mbrec:alloc", "solver workspace allocation failed: %s",
                             hipGetErrorString(e));
        }
        all.push_back(*p);
        return MBREC_OK;
    }
};

// out = A'A v (normalised model), via the sinogram scratch
static int ata(Plan *p, const float *v, float *out, float *ts, int B, hipStream_t s)
{
    int rc = op_forward(p, v, ts, B, true, s);
    if (rc) return rc;
    return op_transpose(p, ts, out, B, true, s);
}

int nnls_l1(Plan *p, const float *y, int B, int coupled, int num_iter, double eta, float *rec, hipStream_t s)
{
    const size_t n = (size_t)p->P, ns = (size_t)p->D * p->nT;
    const int G = coupled ? 1 : B;
    L1Work w;
    float *bp, *aty, *dp, *bin, *r, *rt, *x, *xmin, *uc, *pp, *q, *vh, *uh, *tp, *ts;
    double *part;
    CgsScal *sc;
    int rc;
    float **vecs[] = {&bp, &aty, &dp, &bin, &r, &rt, &x, &xmin, &uc, &pp, &q, &vh, &uh, &tp};
    for (float **v : vecs)
        if ((rc = w.alloc(v, n * B))) return rc;
    if ((rc = w.alloc(&ts, ns * B)) || (rc = w.alloc(&part, (size_t)kMaxQ * B * kRedBlocks)) ||
        (rc = w.alloc(&sc, G)))
        return rc;
    L1Consts k;
    k.lam = 1.0f;                       // lambdaParam = 1 (rec_nn_with_L1_eye_reg.m:25)
    k.lam_eta = (float)(1.0 * eta);     // lambdaParam * etaReg
    k.tol = 1e-8;                       // inner_tol (nnls_L1_eye_matrix_less.m:35)
    k.coupled = coupled;
    float *u = rec;
    // A'y: start value u0 (rec_nn_with_L1_eye_reg.m:32 seeds nnls_L1_eye_matrix_less.m:28) and b_in term
    if ((rc = op_transpose(p, y, aty, B, true, s))) return rc;
    MBREC_HIP(hipMemcpyAsync(u, aty, sizeof(float) * n * B, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, bp, n * B);
    for (int it = 0; it < num_iter; ++it) {
        hipLaunchKernelGGL(k_l1_setup, vgrid(B), dim3(256), 0, s, u, bp, aty, dp, bin, r, rt, x, xmin, n, B, k, part);
        hipLaunchKernelGGL(k_cgs_init, dim3(1), dim3(256), 0, s, sc, part, G, B, k);
        for (int ii = 1; ii <= 2; ++ii) {
            hipLaunchKernelGGL(k_dot, vgrid(B), dim3(256), 0, s, rt, r, n, B, (int)Q_B, part);
            hipLaunchKernelGGL(k_cgs_rho, dim3(1), dim3(256), 0, s, sc, part, G, B, k, ii);
            hipLaunchKernelGGL(k_cgs_dir, vgrid(B), dim3(256), 0, s, r, q, uc, pp, n, B, k, sc, ii);
            if ((rc = ata(p, pp, tp, ts, B, s))) return rc;
            hipLaunchKernelGGL(k_l1_apply, vgrid(B), dim3(256), 0, s, tp, pp, vh, rt, n, B, k, (int)Q_C, part);
            hipLaunchKernelGGL(k_cgs_alpha, dim3(1), dim3(256), 0, s, sc, part, G, B, k);
            hipLaunchKernelGGL(k_cgs_q, vgrid(B), dim3(256), 0, s, uc, vh, q, uh, x, n, B, k, sc, part);
            if ((rc = ata(p, uh, tp, ts, B, s))) return rc;
            hipLaunchKernelGGL(k_cgs_r, vgrid(B), dim3(256), 0, s, tp, uh, r, n, B, k, sc, part);
            hipLaunchKernelGGL(k_cgs_post, dim3(1), dim3(256), 0, s, sc, part, G, B, k);
            hipLaunchKernelGGL(k_cgs_keepmin, vgrid(B), dim3(256), 0, s, x, xmin, n, B, k, sc);
        }
        // non-convergence: the minimal-residual iterate if its true residual is not larger
        if ((rc = ata(p, xmin, tp, ts, B, s))) return rc;
        hipLaunchKernelGGL(k_cgs_rcomp, vgrid(B), dim3(256), 0, s, bin, tp, xmin, n, B, k, part);
        hipLaunchKernelGGL(k_cgs_select, dim3(1), dim3(256), 0, s, sc, part, G, B, k);
        hipLaunchKernelGGL(k_l1_update, vgrid(B), dim3(256), 0, s, x, xmin, dp, u, bp, n, B, k, sc);
    }
    MBREC_HIP(hipGetLastError());
    MBREC_HIP(hipStreamSynchronize(s));
    return MBREC_OK;
}

}  // namespace mbrec

using namespace mbrec;

extern "C" int mbrec_rec_nn_with_L1_eye_reg_dev(mbrec_plan h, const float *data, int B, int coupled,
                                                int num_iter, double regParam, float *rec, mbrec_stream s)
{
    if (!h || !data || !rec || B <= 0 || num_iter < 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "rec_nn_with_L1_eye_reg: invalid arguments");
    return nnls_l1(&h->p, data, B, coupled, num_iter, regParam, rec, (hipStream_t)s);
}

// mbrec_internal.h -- shared definitions of the MI355X MSOT operator library.
//
// Operator formulation (exact, not an approximation):
//   For one detector row y[0..nT) the transpose MEX loop
//   (transposeDelayAndSumMex.cpp:171-200, transposeMex.cpp:277-320) evaluates
//       g(tau) = sum_r y[r] * max(0, 1 - |r - tau| / R)
//   at every pixel's travel time tau.  Each term is a hat in tau with kinks at
//   r - R, r, r + R, so g is piecewise linear with kinks only at the three
//   phases {0, frac(R), 1 - frac(R)} of every sample interval.  Sampling g at
//   those breakpoints ("bins", 3 per sample) therefore represents it exactly,
//   and each (pixel, detector) pair needs one linear interpolation between two
//   adjacent bins instead of ~2R taps.  With an SIR window the pair reads the
//   same bins at tau + s - floor(Ls/2) (same phase, shifted by whole samples).
//   The forward operator is the exact adjoint: each pair deposits its two
//   interpolation weights into bins, and a short polyphase filter turns bins
//   back into samples.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include "../../include/mbrec.h"

#define MBREC_VERSION 1
#define MBREC_MAX_TAPS 64   // per-phase hat taps: 2*floor(R)+3 <= 64  =>  R < 30.5

namespace mbrec {

// Hat weights of the three breakpoint phases, passed by value to kernels.
struct HatTaps {
    float w[3][MBREC_MAX_TAPS];  // w[m][k - kmin] = max(0, 1 - |k - phi_m| / R)
    int kmin;                    // first tap offset (= -floor(R) - 1)
    int nw;                      // taps per phase (= 2*floor(R) + 3)
};

// Decoding of a fixed-point travel time into (bin index, interpolation weight).
struct BinGeom {
    float phi1, phi2;            // sorted interior breakpoint phases in [0,1)
    float iw0, iw1, iw2;         // 1 / segment widths (0 for empty segments)
    int K;                       // bins cover samples n in [-K, nT + K)
    int NB;                      // bins per row = 3 * (nT + 2K) + 1
    int nT;
    int F;                       // fractional bits of the travel-time format
};

struct Plan {
    int Nz = 0, Nx = 0, P = 0, D = 0, nT = 0;
    double R = 0.0;
    int TZ = 8, TX = 8;          // device pixel tile
    int Pd = 0;                  // device pixel slots (tile-padded)
    int32_t *d_pix = nullptr;    // [Pd] slot -> pixel index or -1
    int32_t *d_tq = nullptr;     // [D][Pd] fixed-point travel times
    BinGeom geom{};
    HatTaps taps{};
    // SIR
    int use_sir = 0, Ls = 0, Nsir = 0;
    int32_t *d_sidx = nullptr;   // [D][Pd] 0-based SIR index
    float *d_sir = nullptr;      // [Nsir][Ls]
    int32_t *d_swin = nullptr;   // [Nsir][2] 0-based first tap, tap count
    int sir_maxlen = 0;
    // impulse response
    float *d_ir = nullptr;       // [Lir] or [D][Lir]
    int Lir = 0, ir_per_det = 0;
    // prefilter
    float *d_kz = nullptr, *d_kx = nullptr;
    int Kz = -1, Kx = -1;        // -1 = no prefilter
    double sigma = 1.0;
    // workspace
    int S = 4;                   // pixel splits of the forward scatter
    float *d_bins = nullptr;     // [S][Bcap][D][NB]
    int bins_bcap = 0;
    float *d_img_tmp = nullptr;  // [Bcap][P] x2
    float *d_sino_tmp = nullptr; // [Bcap][D][nT]
    int tmp_bcap = 0;
    size_t table_bytes = 0;
};

// error state -------------------------------------------------------------
int set_error(int code, const char *id, const char *fmt, ...);
int hip_check(hipError_t e, const char *what);
#define MBREC_HIP(call)                                                     \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) return ::mbrec::hip_check(_e, #call);         \
    } while (0)

// plan workspace ----------------------------------------------------------
int ensure_bins(Plan *p, int B);
int ensure_tmp(Plan *p, int B);

// operator launchers (device buffers, planar layouts) ----------------------
int launch_scatter(Plan *p, const float *img, int B, hipStream_t s);           // img -> bins partials
int launch_bins_to_sino(Plan *p, float *sino, int B, bool model, hipStream_t s);
int launch_sino_to_bins(Plan *p, const float *sino, int B, bool model, hipStream_t s);
int launch_gather(Plan *p, float *img, int B, float scale, hipStream_t s);      // bins -> img
int launch_prefilter(Plan *p, const float *in, float *out, float *tmp, int B, bool transp,
                     hipStream_t s);

int time_stage(Plan *p, int which, int B, int reps, hipStream_t s, float *ms);

int op_forward(Plan *p, const float *img, float *sino, int B, bool model, hipStream_t s);
int op_transpose(Plan *p, const float *sino, float *img, int B, bool model, hipStream_t s);

}  // namespace mbrec

struct mbrec_plan_s {
    mbrec::Plan p;
};

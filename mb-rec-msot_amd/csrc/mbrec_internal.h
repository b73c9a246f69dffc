// mbrec_internal.h -- shared definitions of the MI355X MSOT operator library.
//
// Operator formulation (exact, not an approximation):
//   For one detector row y[0..nT) the transpose MEX loop
//   (transposeDelayAndSumMex.cpp:67-96, transposeMex.cpp:115-156) evaluates
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
#include <vector>
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
constexpr int kTileRowsMax = 256;   // tile gathers: longest bin window they stage (rows)
constexpr int kRecPad = 512;   // zero records after the pull table (runs of up to 512 read past the end)
constexpr int kSirWinPad = 32;   // zero weights on each side of a window in the pull form's SIR table
constexpr int kSirBandTS = 30;   // samples per wave of the band-form SIR forward

struct BinGeom {
    float phi1, phi2;            // sorted interior breakpoint phases in [0,1)
    float iw0, iw1, iw2;         // 1 / segment widths (0 for empty segments)
    int K;                       // bins cover samples n in [-K, nT + K)
    int NB;                      // bins per row = 3 * (nT + 2K) + 1
    int nT;
    int F;                       // fractional bits of the travel-time format
};

// launch choices picked by the autotuner for one batch stride
struct TuneChoice {
    int Bp = 0;          // batch stride the choices were timed for (0 = none)
    int gather = -1;     // gather variant index
    int scatter_bc = 0;  // scatter LDS batch chunk
};

// tile forward (one sinogram): rows of a (block, detector) window the kernel stages in LDS;
// a plan whose windows are longer gets no tile form
constexpr int kTfMaxRows = 2048;

struct Plan {
    int Nz = 0, Nx = 0, P = 0, D = 0, nT = 0;
    double R = 0.0;
    int TZ = 16, TX = 16;        // device pixel tile (256 slots; 64-slot runs are 8 x 8 squares)
    int Pd = 0;                  // device pixel slots (tile-padded)
    int32_t *d_pix = nullptr;    // [Pd] slot -> pixel index or -1
    int32_t *d_tq = nullptr;     // [D][Pd] fixed-point travel times (transpose gather)
    int2 *d_ptab = nullptr;      // [Pd][D] decoded pairs (bin row d*NB + j or the zero row D*NB, u bits)
    int2 *d_twin = nullptr;      // [Pd / 256][D] bin window (first bin, length) of each 256-slot tile
    int tile_wmax = 0;           // longest tile window
    int2 *d_ttab = nullptr;      // [Pd / 256][D][256] (bin row inside the tile window, or tile_zr; u bits)
    int tile_zr = 0;             // first of the two zero rows after a staged tile window (64 * ceil(wmax / 64))
    int32_t *d_tord = nullptr;   // [tile_launch] tile of each workgroup of the batch tile gather, or -1
                                 // (XCD-aware order; null: workgroup = tile)
    int tile_launch = 0;         // workgroups per batch chunk and detector group of that gather
    uint32_t *d_tq1 = nullptr;   // [Pd / 256][D][256] packed (row inside the tile window, or tile_zr: 9 bits |
                                 // u in units of 2^-23: 23 bits); tile-major stream of the one-sinogram gather
    BinGeom geom{};
    HatTaps taps{};
    // forward scatter: per detector, (pixel, travel time) records grouped into buckets
    // of Wb consecutive bins; bucket k of detector d deposits into bins
    // [k*Wb + lo_off, k*Wb + lo_off + WL)
    // Without SIR (pull = 1) the forward is gather-only instead: d_rec holds
    // (pixel, weight u as float bits) sorted by bin and d_rec_off [D][NB + 1]
    // CSR offsets (see build_pull_records).
    int pull = 0;
    int Wb = 384, nbk = 0, lo_off = 0, hi_off = 1, WL = 0;
    int2 *d_rec = nullptr;       // [records] (pixel, fixed-point travel time) / (pixel, u)
    int2 *d_rec1 = nullptr;      // [records] (slot, u): the pull records with slot indices (one sinogram)
    // tile forward for one sinogram: blocks of 4 x 4 device tiles (64 x 64 pixels, 4096 slots);
    // per (block T, detector d) the pairs sorted by bin j, each a 32-bit record (slot in the
    // block: 12 bits | u in units of 2^-20: 20 bits), row offsets (uint16, relative to the
    // record start) over the window [jmin, jmin + nrow), and a window partial of nrow + 1 bins
    uint32_t *d_tfrec = nullptr;
    uint16_t *d_tfroff = nullptr;
    int4 *d_tfdesc = nullptr;    // [nblocks][D] + 1 (record start, offset / partial start, jmin, nrow);
                                 // the extra entry's record start = the total (run lengths by difference)
    float *d_tfpart = nullptr;   // partial windows (one sinogram), summed per bin by a second pass
    // the same records with each pair's rows in order of decreasing record count (a row's
    // records keep their order), and per sorted position (first record | row << 16), entry
    // nrow = (record count, 0); indexed like d_tfrec / d_tfroff
    uint32_t *d_tfsrec = nullptr;
    uint32_t *d_tfsro = nullptr;
    int tf_nbz = 0, tf_nbx = 0;  // blocks along z / x
    size_t tf_bytes = 0;
    int32_t *d_rec_off = nullptr;   // [D][nbk + 1] record offsets / [D][NB + 1]
    int32_t *d_rec_k = nullptr;  // [records] 0-based SIR index (SIR plans)
    // SIR forward in pull form: one record per (pixel, detector) pair with a tap in the
    // bin range, int4 (pixel | phase << 28, first affected sample nf = n + s0 - Ls/2 + K as
    // 16-bit two's complement | tap count << 16, u bits, d_sirw offset of the first tap),
    // sorted by (detector, phase, nf);
    // CSR over nf in [sp_lo, sp_lo + sp_n) per (detector, phase): d_srec_off[(3 d + m) (sp_n + 1) + i]
    int4 *d_srec = nullptr;
    float *d_sirw = nullptr;     // [Nsir][maxlen + 2 kSirWinPad]: each SIR window between zero margins
    // SIR transpose in LDS tile form: per (256-slot tile, detector) the first bin row of the
    // staged window and per pair int4 (row of its first in-range tap inside the window,
    // u bits, d_sirw offset of that tap's weight, in-range tap count)
    int32_t *d_stwin = nullptr;  // [Pd / 256][D]
    int4 *d_stab = nullptr;      // [Pd / 256][D][256]
    int stile_zr = 0;            // staged rows: 64 * ceil(longest window / 64)
    int stile_wmax = 0;          // longest window
    size_t stab_bytes = 0;
    int32_t *d_srec_off = nullptr;
    int sp_lo = 0, sp_n = 0, Nsir_build = 0;
    size_t srec_bytes = 0;
    size_t nrec = 0;
    size_t rec_bytes = 0;              // device bytes of d_rec / d_rec_off / d_rec_k
    // SIR
    int use_sir = 0, Ls = 0, Nsir = 0;
    int32_t *d_sidx = nullptr;   // [D][Pd] 0-based SIR index
    float *d_sir = nullptr;      // [Nsir][Ls]
    int32_t *d_swin = nullptr;   // [Nsir][2] 0-based first tap, tap count
    int sir_maxlen = 0;
    // impulse response (taps of the convolution and of the correlation = flipped IR)
    float *d_ir = nullptr;       // [Lir] or [D][Lir]
    float *d_irf = nullptr;      // flipped
    double *d_ir64 = nullptr;    // [rows][16 ceil(Lir / 16)] the same taps in FP64, zero-padded (FIR kernel)
    double *d_irf64 = nullptr;   // flipped
    int Lir = 0, ir_per_det = 0;
    int fir_mfma = 0;            // IR filter: 1 Toeplitz products on the FP64 matrix cores (one IR for all detectors), 0 VALU
    // prefilter
    float *d_kz = nullptr, *d_kx = nullptr;
    int Kz = -1, Kx = -1;        // -1 = no prefilter
    double sigma = 1.0;
    // workspace, sized for bcap sinograms
    int bcap = 0;
    float *d_bins = nullptr;     // [D][NB][Bp]      transpose bins
    float *d_hpart = nullptr;    // [D][nbk][WL][Bp] forward bucket windows
    float *d_imgI = nullptr;     // [Pd..P][Bp]      interleaved image (scatter input)
    float *d_img_tmp = nullptr;  // 2 x [B][P]       prefilter temporaries
    float *d_sino_tmp = nullptr; // [B][D][nT]       samples before/after the IR filter
    float *d_gpart = nullptr;    // [max(8 B, 16)][P]  partial images of the detector-grouped gathers
    float *d_img_slot = nullptr; // [Pd]  one sinogram's image in slot order (pull forward)
    float *d_sband = nullptr;    // [D][band waves][3 kSirBandTS + 1][Bp]  band-form SIR forward rows (B > 1)
    int sband_nwave = 0, sband_nfb = 0;
    size_t table_bytes = 0;
    size_t work_bytes = 0;
    int pins = 0;                // batches whose captured graphs use this plan's buffers
    TuneChoice tune;
};

static inline int batch_stride(int B) { return B <= 1 ? 1 : (B + 3) / 4 * 4; }

// q = e / d, r = e % d for 0 <= e < 2^24, 0 < d < 2^16 (float reciprocal, one correction step)
__device__ __forceinline__ void divmod_small(int e, int d, float inv, int &q, int &r)
{
    q = (int)((float)e * inv);
    r = e - q * d;
    if (r < 0) { --q; r += d; }
    else if (r >= d) { ++q; r -= d; }
}

// error state -------------------------------------------------------------
int set_error(int code, const char *id, const char *fmt, ...);
int hip_check(hipError_t e, const char *what);
#define MBREC_HIP(call)                                                     \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) return ::mbrec::hip_check(_e, #call);         \
    } while (0)

// plan workspace ----------------------------------------------------------
int ensure_work(Plan *p, int B);
int build_records(Plan *p, const std::vector<int32_t> &slot_pix, const std::vector<int32_t> &tq,
                  const int32_t *sidx, const int32_t *win, int Ls);
void host_decode(int32_t q, const BinGeom &g, int &nn, int &m);

int time_stage(Plan *p, int which, int B, int reps, hipStream_t s, float *ms);
int autotune(Plan *p, int B, int reps, hipStream_t s, int *gather_out, int *scatter_bc_out);
int set_launch(Plan *p, int B, int gather, int scatter);
int num_gather_variants();

// win (device, [B], optional): per-sinogram sample window [x, y); samples outside
// it are zero in the output (forward) and ignored in the input (transpose), both
// before and after the IR filter -- a sinogram cropped to that window.
int op_forward(Plan *p, const float *img, float *sino, int B, bool model, hipStream_t s,
               const int2 *win = nullptr);
int op_transpose(Plan *p, const float *sino, float *img, int B, bool model, hipStream_t s,
                 const int2 *win = nullptr);
// op_transpose on a workspace that is already in place (p->d_sino_tmp, d_bins with its two
// zero rows, d_img_tmp of 2 * bcap * P, d_gpart): launches kernels only
int op_transpose_run(Plan *p, const float *sino, float *img, int B, bool model, hipStream_t s,
                     const int2 *win = nullptr);

}  // namespace mbrec

struct mbrec_plan_s {
    mbrec::Plan p;
};

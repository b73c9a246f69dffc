/*
 * mbrec.h -- C ABI of the MI355X-native model-based MSOT operators and
 * non-negative least-squares solvers.
 *
 * Two layers:
 *
 * 1. MEX drop-ins.  One entry point per reference MEX file, same argument
 *    order, dtypes and output shapes (FP64 host buffers in, FP64 host buffer
 *    out, column-major MATLAB layouts).  A MATLAB gateway (mexFunction) only
 *    unpacks mxArrays and calls these; it raises the same error IDs as the
 *    reference when a type check fails.  Device-resident tables are cached by
 *    (pointer, dims, content hash), so tables are uploaded once.
 *
 * 2. Device API.  A plan holds the device-resident geometry (travel times,
 *    SIR tables, impulse response, prefilter) and runs the operators and
 *    solvers on device buffers (FP32 bulk, FP64 scalars) on a caller stream.
 *
 * All functions return MBREC_OK (0) or a negative status; mbrec_last_error()
 * and mbrec_last_error_id() describe the last failure of the calling thread.
 *
 * Layouts.  MEX image [W x P] (wavelength fastest), MEX sinogram [W x nT x D],
 * travel times / SIR indices [P x D] (pixel fastest, tt[d*P + x]), SIR table
 * [Ls x Nsir], SIR windows [2 x Nsir] (1-based, inclusive).  Device images are
 * planar [B][P] with pixel x = iz + Nz*ix, device sinograms planar [B][D][nT]
 * (the MATLAB vector layout of MSOTmodel.m:194,202,226).
 */
#ifndef MBREC_H
#define MBREC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBREC_OK 0
#define MBREC_ERR_ARG (-1)      /* invalid argument / shape */
#define MBREC_ERR_TYPE (-2)     /* wrong input class (MEX type checks) */
#define MBREC_ERR_HIP (-3)      /* HIP runtime error */
#define MBREC_ERR_NODEV (-4)    /* no usable GPU */
#define MBREC_ERR_RANGE (-5)    /* travel times outside the fixed-point range */

typedef struct mbrec_plan_s *mbrec_plan;  /* opaque device-resident geometry */
typedef void *mbrec_stream;                /* hipStream_t; NULL = null stream */

/* ---- library ------------------------------------------------------------ */
const char *mbrec_last_error(void);
const char *mbrec_last_error_id(void);
int mbrec_version(void);
/* Selects and initialises a device (hipSetDevice); 0 on success. */
int mbrec_init(int device);

/* ---- 1. MEX drop-ins ----------------------------------------------------- */

/* forwardMex (reconstruction_model_based/models/forwardMex.cpp:16-162):
 * sig = forwardMex(image[W x P], sirIndices[P x D] int32, travelTimes[P x D],
 *                  sir[Ls x Nsir], numTimesteps int32, sirTimestepsStartAndEnd[2 x Nsir] int32, R)
 * out: [W x nT x D], overwritten (the MEX zero-fills, :84). */
int mbrec_forwardMex(const double *image, int W, int P, const int32_t *sirIndices,
                     const double *travelTimes, int D, const double *sir, int Ls, int Nsir,
                     int32_t numTimesteps, const int32_t *sirTimestepsStartAndEnd,
                     double R, double *out);

/* transposeMex (reconstruction_model_based/models/transposeMex.cpp:16-158):
 * img = transposeMex(sigMat[W x nT x D], sirIndices, travelTimes, sir, sirSE, R)
 * out: [W x P]. */
int mbrec_transposeMex(const double *sigMat, int W, int nT, int D, const int32_t *sirIndices,
                       const double *travelTimes, int P, const double *sir, int Ls, int Nsir,
                       const int32_t *sirTimestepsStartAndEnd, double R, double *out);

/* forwardDelayAndSumMex (reconstruction_model_based/models/forwardDelayAndSumMex.cpp:15-104):
 * sig = forwardDelayAndSumMex(image[W x P], travelTimes[P x D], numTimesteps int32, R)
 * Samples outside [0, nT) are dropped (the reference writes out of bounds, :80-99). */
int mbrec_forwardDelayAndSumMex(const double *image, int W, int P, const double *travelTimes,
                                int D, int32_t numTimesteps, double R, double *out);

/* transposeDelayAndSumMex (reconstruction_model_based/models/transposeDelayAndSumMex.cpp:15-98):
 * img = transposeDelayAndSumMex(sigMat[W x nT x D], travelTimes[P x D], R) */
int mbrec_transposeDelayAndSumMex(const double *sigMat, int W, int nT, int D,
                                  const double *travelTimes, int P, double R, double *out);

/* Type-check helper shared by the MEX gateways: maps an argument position and
 * an observed class to the reference error ID ("forwardMex:notDouble" /
 * "forwardMex:notInt32", forwardMex.cpp:30-50) and sets the last error. */
int mbrec_mex_type_error(const char *mex_name, int arg_index, int expect_int32);

/* Drops every cached device table (the gateways call this from mexAtExit). */
void mbrec_mex_cache_clear(void);

/* ---- 2. Device API -------------------------------------------------------- */

/* Geometry: travelTimes [P x D] FP64 host (0-based fractional samples, as
 * MSOTmodel.m:380-381 produces them).  Nz, Nx give the pixel grid (P = Nz*Nx)
 * and enable the tiled device pixel order; pass Nz = P, Nx = 1 when unknown.
 * Replaces the per-call table marshalling of MSOTmodel.m:195-252. */
int mbrec_plan_create(mbrec_plan *plan, int Nz, int Nx, int D, int nT, double R,
                      const double *travelTimes);
/* SIR tables as forwardMex takes them (MSOTmodel.m:196-200). */
int mbrec_plan_set_sir(mbrec_plan plan, const int32_t *sirIndices, const double *sir, int Ls,
                       int Nsir, const int32_t *sirTimestepsStartAndEnd);
/* Impulse response IR [Lir] or [Lir x D] (per_detector), MSOTmodel.m:259-269. */
int mbrec_plan_set_ir(mbrec_plan plan, const double *ir, int Lir, int per_detector);
/* Separable circular prefilter (Discretization.m:54-72): taps kz[0..2Kz],
 * kx[0..2Kx] for offsets -K..K of the kernel k[n] = ifft(1/phi_hat)[n], K <= N/2 per
 * axis.  With K = N/2 (N even) offsets -K and +K reach the same sample: the caller puts
 * the kernel's N/2 coefficient in one of them and 0 in the other. */
int mbrec_plan_set_prefilter(mbrec_plan plan, const double *kz, int Kz, const double *kx, int Kx);
/* model_normalization_factor (MSOTmodel.m:271-287). */
int mbrec_plan_set_sigma(mbrec_plan plan, double sigma);
int mbrec_plan_destroy(mbrec_plan plan);
/* Bytes the plan holds on the device (tables + workspace). */
size_t mbrec_plan_device_bytes(mbrec_plan plan);
/* Fixed-point travel-time format: number of fractional bits. */
int mbrec_plan_tau_frac_bits(mbrec_plan plan);

/* Autotuner: times every valid launch variant of the transpose gather (sinograms
 * per pass x detector unroll) and of the forward scatter (LDS batch chunk) for a
 * batch of B on this plan (reps launches each, HIP events on stream s, after >= 20 ms
 * of untimed work so that the GPU's clocks have ramped up), times the three fastest
 * of each again in two alternating rounds, and keeps the fastest for later calls with
 * the same batch stride.  choice (may be NULL) receives {gather variant, scatter
 * chunk}.  Results do not depend on the choice beyond FP32 summation order. */
int mbrec_plan_autotune(mbrec_plan plan, int B, int reps, mbrec_stream s, int *choice);
/* Set the launch choice for batches of B by hand: gather variant in
 * [0, mbrec_num_gather_variants()) or -1 (automatic); forward choice 0:
 * automatic.  Plans with SIR: 1000 the pull form (one sinogram: lane groups of
 * samples, the default there; batches: 24 samples x 2 records in flight per wave), 1001 / 1002 / 1003
 * batch pull forms with (16, 4), (16, 2), (32, 1), 1004 the band form for batches
 * (each record once, overlapping rows summed in a second pass; the default for batches);
 * bucketed scatter: > 0 the
 * record-lane kernel with that LDS batch chunk (a multiple of 4), < 0 the
 * batch-lane kernel with lane groups of -value (16, 32 or 64) sinograms.  Plans
 * without SIR (pull form): 1, 4, 8, 16 bins per thread pass of that many
 * sinograms; 2, 3 record streaming for one sinogram (1 or 4 bins per lane); 5, 6, 7 the
 * tile form for one sinogram with 8, 4, 2 detectors per workgroup (4 x 4-tile blocks,
 * 32-bit records staged in LDS, window partials summed per bin); 9, 10 its lean form
 * (scalar descriptors, unconditional loads, sums of v and t v per row; 2, 4 detectors
 * per workgroup); 11, 12, 13 the lean form over rows sorted by record count (the same
 * sums; 2, 2, 4 detectors per workgroup, records read in chunks of 4, 2, 2; 11 is the
 * default for one sinogram);
 * -(LB + 100 mode) batch-lane with lane groups LB and mode 0 (plain order), 1
 * (XCD-aware order), 2 (XCD-aware, 16 pairs in flight); -364 / -464 / -564 /
 * -664 / -764 / -864 wave-uniform, -964 / -1064 / -1164 wave-uniform with raw
 * buffer loads (bins per wave and rows in flight differ).  Invalid choices are
 * rejected (mbrec:args). */
int mbrec_plan_set_launch(mbrec_plan plan, int B, int gather_variant, int scatter_choice);
/* Kernel of the IR convolution / correlation stage (MSOTmodel.m:259-269, FFTs in FP64
 * there): mfma = 0 (default) the register-blocked vector kernel, mfma = 1 Toeplitz
 * products on the FP64 matrix cores when one IR serves every detector.  Both sum the
 * exact products of the FP32 samples and taps in FP64; the results differ by summation
 * order only.  No counterpart in the reference: a speed choice. */
int mbrec_plan_set_fir_kernel(mbrec_plan plan, int mfma);
int mbrec_num_gather_variants(void);

/* Profiling: average duration (ms, HIP events on stream s) of one operator
 * stage over reps launches for a batch of B: 0 = forward pair scatter into
 * bucketed bin windows, 1 = bins -> samples (window sum + hat filter),
 * 2 = IR convolution (FIR), 3 = samples -> bins (hat filter), 4 = transpose
 * pair gather.  Inputs are zero buffers; timing does not depend on their values. */
int mbrec_plan_time_stage(mbrec_plan plan, int stage, int B, int reps, mbrec_stream s, float *ms);

/* Bare MEX loop (no prefilter, IR or 1/sigma), batch of B planar images
 * [B][P] -> sinograms [B][D][nT] and back. */
int mbrec_mex_forward_dev(mbrec_plan plan, const float *img, float *sino, int B, mbrec_stream s);
int mbrec_mex_transpose_dev(mbrec_plan plan, const float *sino, float *img, int B, mbrec_stream s);

/* Normalised model operators model.Funcs.applyForward / applyTranspose
 * (MSOTmodel.m:285-286): prefilter, MEX loop, IR convolution, 1/sigma. */
int mbrec_model_forward_dev(mbrec_plan plan, const float *img, float *sino, int B, mbrec_stream s);
int mbrec_model_transpose_dev(mbrec_plan plan, const float *sino, float *img, int B, mbrec_stream s);
/* The same on cropped acquisitions: sinogram b is the window (first, count) =
 * windows[2b], windows[2b+1] (host) of the plan's sample axis, i.e. the model of
 * MSOTmodel.m with numCroppedSamplesAtSinogramStart increased by first and
 * count samples; sino keeps the plan's [B][D][nT] layout, zero (forward) or
 * ignored (transpose) outside each window.  Synchronous. */
int mbrec_model_forward_window_dev(mbrec_plan plan, const float *img, float *sino, int B, const int32_t *windows,
                                   mbrec_stream s);
int mbrec_model_transpose_window_dev(mbrec_plan plan, const float *sino, float *img, int B,
                                     const int32_t *windows, mbrec_stream s);

/* Model transpose that allocates nothing (no counterpart in the reference): for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same kernels and result as mbrec_model_transpose_dev;
 * its scratch (IR-filtered sinograms, bins, prefilter images, detector-group partial
 * images) is a file-scope device array of 64 MiB per GPU, held until the process ends.
 * A request that needs more (large B or grids) fails with mbrec:scratchTooSmall.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU race. */
int mbrec_model_transpose_noalloc_dev(mbrec_plan plan, const float *sino, float *img, int B, mbrec_stream s);

/* rec_nn_with_L2_reg (reconstruction/rec_nn_with_L2_reg.m:1-66) with RegL2 = identity
 * (mbrec_rec_nn_with_L2_reg_terms_dev below takes other L2 terms)
 * and nnls_L2_less (nnls_L2_less.m:1-101): 50-style outer splitting, 2 LSQR inner
 * iterations, lambda = 1, beta = 1e-3.  data [B][D][nT], rec [B][P].
 * coupled != 0 solves the B sinograms as one stacked problem with shared LSQR
 * scalars (the reference's multi-wavelength call); coupled == 0 solves each
 * independently.  init (nullable) seeds the outer iterate w0 (warm start; the
 * reference ignores its initial image, nnls_L2_less.m:28-29).  lcurve (host,
 * nullable) receives per group {||A rec - y||^2, ||rec||^2}. */
int mbrec_rec_nn_with_L2_reg_dev(mbrec_plan plan, const float *data, int B, int coupled,
                                 int num_iter, double lambdaL2, const float *init,
                                 float *rec, double *lcurve, mbrec_stream s);

/* rec_nn_with_L2_reg with general L2 terms (rec_nn_with_L2_reg.m:1-66,
 * incorporate_l2_reg_into_model_functions.m:1-21, combine_forward_and_transpose_with_reg.m:6-14,
 * combine_forward_and_transpose_with_two_regs.m:8-14): objective ||A x - y||^2 +
 * lambdaL2 ||R1 x||^2 + lambdaL2_2 ||R2 x||^2 over x >= 0, the stacked operator
 * [A; sqrt(lambdaL2) R1; sqrt(lambdaL2_2) R2] inside nnls_L2_less.  terms[0] = (lambdaL2,
 * RegL2), terms[1] = (lambdaL2_2, RegL2_2); as in the reference the second term enters
 * only when lambdaL2 > 0.  Each term's operator maps one image [P] to one image [P], per
 * sinogram: MBREC_REG_IDENTITY (RegL2 = @(x) x), MBREC_REG_DEL2 (MATLAB del2(img, 1) per
 * image, used as its own transpose like the reference's laplacian_per_wavelength,
 * reconstruction_overview_script.m:50-51), or MBREC_REG_CALLBACK: apply(ctx, in, out, B, P,
 * transp, s) writes out = R in (transp = 0) or R' in (transp = 1) for B images [B][P] on
 * stream s, enqueuing work only (it may then be captured) and returning 0.  lcurve (host,
 * nullable) receives per group {||A rec - y||^2, ||R1 rec||^2, ||R2 rec||^2} (0 for an
 * absent term).  Other arguments as mbrec_rec_nn_with_L2_reg_dev. */
#define MBREC_REG_IDENTITY 0
#define MBREC_REG_DEL2 1
#define MBREC_REG_CALLBACK 2
typedef int (*mbrec_reg_apply)(void *ctx, const float *in, float *out, int B, int P, int transp, mbrec_stream s);
typedef struct {
    double lambda;
    int kind;
    mbrec_reg_apply apply;
    void *ctx;
} mbrec_l2_term;
int mbrec_rec_nn_with_L2_reg_terms_dev(mbrec_plan plan, const float *data, int B, int coupled, int num_iter,
                                       const mbrec_l2_term *terms, int nterms, const float *init, float *rec,
                                       double *lcurve, mbrec_stream s);

/* laplacian_per_wavelength (reconstruction/laplacian_per_wavelength.m:1-10): MATLAB
 * del2(img, 1) of B device images [B][Nx][Nz] (z fastest), out = in not allowed. */
int mbrec_del2_dev(const float *in, float *out, int Nz, int Nx, int B, mbrec_stream s);

/* Independent solves of B sinograms with a lambdaL2 of their own (host array
 * [B]), as B separate rec_nn_with_L2_reg.m calls would use; otherwise as
 * mbrec_rec_nn_with_L2_reg_dev with coupled = 0 (lcurve: 2 values per sinogram).
 * init / init_dual (may be NULL): starting outer iterate w and scaled dual v of
 * nnls_L2_less.m:28-29 (zero there) -- the warm start of a learned estimate.
 * final_dual (may be NULL) receives the final scaled dual v [B][P]: (rec, final_dual)
 * is the loop's whole state, so passing them back as init / init_dual continues the
 * same iteration sequence. */
int mbrec_rec_nn_with_L2_reg_batch_dev(mbrec_plan plan, const float *data, int B, int num_iter,
                                       const double *lambdaL2, const float *init, const float *init_dual,
                                       float *rec, float *final_dual, double *lcurve, mbrec_stream s);
/* rec_nn_with_L1_eye_reg (reconstruction/rec_nn_with_L1_eye_reg.m:1-50) and
 * nnls_L1_eye_matrix_less (nnls_L1_eye_matrix_less.m:1-67): split Bregman with
 * 2 CGS iterations of (A'A + I) per outer iteration, lambdaParam = 1, etaReg =
 * regParam, started from A'y.  data [B][D][nT], rec [B][P]; coupled as above. */
int mbrec_rec_nn_with_L1_eye_reg_dev(mbrec_plan plan, const float *data, int B, int coupled,
                                     int num_iter, double regParam, float *rec, mbrec_stream s);

/* rec_nn_with_TV_reg (reconstruction/rec_nn_with_TV_reg.m:1-55) and
 * nnls_sparsa_based_reg_matrix_less (nnls_sparsa_based_reg_matrix_less.m:1-86): the
 * outer splitting with SpaRSA inner solves, Psi = split-Bregman isotropic-TV prox
 * (SB_ITV2.m), Phi = isotropic TV (ITV.m).  inner_iterations = SpaRSA's MaxiterA (2 in
 * the reference, :64), inner_min_iterations = its MiniterA (not passed by the reference,
 * so SpaRSA's default 5 applies there): SpaRSA takes no fewer than MiniterA steps and,
 * past MiniterA, stops after MaxiterA or when the relative objective change is <= 1e-8.
 * Square grids only.  lcurve (host, nullable) receives per group {||A rec - y||^2, Phi(rec)}. */
int mbrec_rec_nn_with_TV_reg_dev(mbrec_plan plan, const float *data, int B, int coupled, int num_iter,
                                 double regParam, int inner_iterations, int inner_min_iterations, float *rec,
                                 double *lcurve, mbrec_stream s);

/* SB_ITV2 (reconstruction/SB_ITV2.m:1-40): u = arg min 1/2||u - g||^2 + mu ITV_B(u) by split
 * Bregman for B square N x N images (device, FP32 in/out, [B][N][N] with z fastest), one
 * mu per image (host array).  The inner cgs(I + B'B, ., 1e-5, 100) of SB_ITV2.m:31 is the
 * exact DCT Neumann solve; N <= 2048.  iterations (host, nullable) = split-Bregman steps
 * taken by the slowest image. */
int mbrec_sb_itv2_dev(const float *g, int N, int B, const double *mu, float *u, int *iterations, mbrec_stream s);

/* ---- 3. Secondary reconstruction paths (FP64) ------------------------------ */

/* dmas_cf (reconstruction_backprojection/dmas_cf.m:19-107): delay-multiply-and-sum
 * with coherence factor.  xpix/zpix: device pixel coordinates [P] (fov_x(:), fov_z(:));
 * sig: device FP32 sinograms [W][D][nT] (MATLAB [nT x D x W]); out: device FP64
 * [4][W][P] = {dmas_cf, dmas, cf, das}.  An arrival sample outside [1, nT] is an
 * error ("dmas_cf:indexOutOfRange"), as MATLAB's indexing would be. */
int mbrec_dmas_cf_dev(const double *xpix, const double *zpix, int P, const float *sig, int W, int D,
                      int nT, double c, double fs, int cropped, double coverage_deg, double radius,
                      double *out, mbrec_stream s);

/* backproject_waveeq (reconstruction_backprojection/backproject_waveeq.m:1-50).
 * xv [Nx], zv [Nz]: host grid vectors of the meshgrid fov_x / fov_y; sig: device FP64
 * [W][D][nT]; psf: host [pm x pn] (time x detector); K: device [Nt x Nt] column-major
 * with Nt >= nT + cropped (bp_kernel output); out: device FP64 [W][Nx][Nz]. */
int mbrec_backproject_waveeq_dev(const double *xv, const double *zv, int Nz, int Nx, const double *sig,
                                 int W, int D, int nT, double c, double fs, int cropped,
                                 double coverage_deg, double radius, const double *psf, int pm, int pn,
                                 const double *K, int Nt, double *out, mbrec_stream s);

/* bp_kernel (reconstruction_backprojection/bp_kernel.m:1-12): K = -Im(H diag(w lambda) H^H)
 * with H = H0^(1)(t_j lambda_l), trapezoid weights w.  t: host [Nt]; K: device [Nt x Nt]. */
int mbrec_bp_kernel_dev(const double *t, int Nt, double *K, mbrec_stream s);

/* upsample_detectors_in_sinogram (load_and_preprocess/upsample_detectors_in_sinogram.m:1-13)
 * as a linear map: out_b [nT x Dup] = sig_b [nT x D] * U', U [Dup x D] column-major
 * (device, the not-a-knot spline map), for nbatch sinograms [nbatch][D][nT]. */
int mbrec_upsample_detectors_dev(const double *U, int Dup, int D, const double *sig, int nT, int nbatch,
                                 double *out, mbrec_stream s);

/* ---- 4. Geometry tables ---------------------------------------------------- */

/* calculate_travel_time_indices (models/MSOTmodel.m:350-382) with calculate_travel_time /
 * refraction (basic_tools/refraction.m:10-24) and fermat_principle (fermat_principle.m:12-32):
 * xpix, zpix device [P] pixel coordinates; xdet, zdet host [D]; membrane_z (host [nm], or
 * NULL for a single speed of sound): the membrane z = m(x) sampled at membrane_x0 + i
 * membrane_dx (linear in between, clamped).  A pixel with z <= m(x) (or with no membrane)
 * travels straight at c_above; otherwise the time is minimised over the membrane crossing
 * x in [-radius, radius] (c_below below, c_above above).  tau (device [D][P]) = t fs - delay
 * - ncrop; x_intersections (device [D][P], may be NULL): the crossing x, NaN for straight
 * paths.  FP64. */
int mbrec_travel_times_dev(const double *xpix, const double *zpix, int P, const double *xdet, const double *zdet,
                           int D, double c_below, double c_above, double radius, const double *membrane_z, int nm,
                           double membrane_x0, double membrane_dx, double fs, double delay, double ncrop,
                           double *tau, double *x_intersections, mbrec_stream s);

/* get_SPMR_and_SPMR_indices (MSOTmodel.m:463-480) and the polar grid
 * (map_grid_point_coordinates_to_polar_coordinates_for_SPMR, :573-597;
 * calculate_required_SPMR_distances_and_angels, :599-616): xv, zv device pixel coordinates
 * [P], or per detector [D][P] (per_detector = 1: the virtual positions of the dual speed of
 * sound); xdet, zdet host [D].  grid (host [7]) receives {dmin, dmax, distance step,
 * largest angle (deg), angle step (deg), number of distances, number of angles}; sidx
 * (device [D][P], may be NULL) the 1-based indices (round(1 + (dist - dmin)/dd) - 1)
 * n_angle + 1 + round(angle_deg / da). */
int mbrec_spmr_grid_dev(const double *xv, const double *zv, int per_detector, int P, const double *xdet,
                        const double *zdet, int D, double resolution_fov, double c, double *grid, int32_t *sidx,
                        mbrec_stream s);
/* The indices alone for a given grid (a precomputed table's, MSOTmodel.m:470-478). */
int mbrec_spmr_index_dev(const double *xv, const double *zv, int per_detector, int P, const double *xdet,
                         const double *zdet, int D, double dmin, double dd, double da, int n_angle, int32_t *sidx,
                         mbrec_stream s);
/* The polar SPMR table on the distance x angle grid (host dists [n_dist], angles_deg
 * [n_angle]), replacing calculate_polarSPMR.m's Field II computation: the element's width
 * split into n_sub (<= 32) point sub-elements, each arrival-time offset deposited as a
 * linear hat on the Ls-sample grid (centre Ls/2), convolved with the pixel-model N-shape
 * taps nshape[Ln] (tap j at sample offset j + nshape_offset).  table device [n_dist *
 * n_angle][Ls] (angle fastest = MATLAB SPMR [Ls x N_angle*N_dist]); per entry, samples
 * below trim x the entry's largest magnitude are set to zero (calculate_polarSPMR.m:152-161,
 * trim 1e-4 there) and windows (device [n_dist * n_angle][2] int32) receive the 1-based
 * first / last non-zero sample (an empty window is (1, 0)). */
int mbrec_spmr_table_dev(const double *dists, int n_dist, const double *angles_deg, int n_angle, double element_width,
                         int n_sub, double c, double fs, const double *nshape, int Ln, int nshape_offset, int Ls,
                         double trim, double *table, int32_t *windows, mbrec_stream s);

/* ---- 5. Preprocessing (load_and_preprocess/, FP64) ------------------------- */

/* filter_butter_zero_phase (load_and_preprocess/filter_butter_zero_phase.m:1-53, FFT
 * branch): sig, out: device FP64 [ncols][nT] (time fastest, MATLAB [nT x cols]); demean
 * along time when fmin > 0; spectrum times |H|^2 of the 8th-order Butterworth low-pass at
 * fmax (if > 0) and 2nd-order high-pass at fmin (if > 0) at the nT bins of the whole
 * circle, as a circular convolution.  out must not alias sig. */
int mbrec_filter_butter_zero_phase_dev(const double *sig, int nT, int ncols, double fs, double fmin, double fmax,
                                       double *out, mbrec_stream s);
/* apply_butterworth_window_to_sinogram (apply_butterworth_window_to_sinogram.m:1-10), in
 * place on device FP64 [ncols][nT]: s = mean + (s - mean) |h(pi t / nT)|, h the order
 * n_butter digital Butterworth band-pass [low_cutoff, high_cutoff] / nT. */
int mbrec_apply_butterworth_window_dev(double *sig, int nT, int ncols, int n_butter, double low_cutoff,
                                       double high_cutoff, mbrec_stream s);
/* interpolate_signals_of_broken_transducers (interpolate_signals_of_broken_transducers.m:1-20):
 * sig, out device FP64 [nbatch][D][nT]; broken (host, 1-based, nbroken entries): each
 * detector's signal becomes the not-a-knot spline across the working detectors evaluated
 * at its index (interp2 'spline'), per time sample.  out must not alias sig. */
int mbrec_interpolate_broken_transducers_dev(const double *sig, int D, int nT, int nbatch, const int32_t *broken,
                                             int nbroken, double *out, mbrec_stream s);

/* ---- batched solves over several geometries, one graph launch -------------
 * No counterpart in the reference (it reconstructs one acquisition per call).
 * ngroups groups, group g = a plan (its own grid and sample axis of nT_g
 * samples) and group_B[g] sinograms: data[g] [B_g][D][nT_g], rec[g] [B_g][P_g],
 * init[g] / init_dual[g] (may be NULL, or the whole array NULL) [B_g][P_g]: warm
 * start of the outer iterate and of the scaled dual -- device buffers
 * owned by the caller, whose addresses are captured.  windows [sum B][2]:
 * (first sample, sample count) of every sinogram on its group's axis -- a
 * sinogram cropped at its start and/or end; samples outside are ignored.
 * lambdaL2 [sum B] (host).  Each sinogram is an independent
 * rec_nn_with_L2_reg.m solve (RegL2 = identity) of num_iter outer iterations.
 * Groups must use distinct plans.  mbrec_batch_run(use_graph = 1) captures all
 * groups' solves into one HIP graph on first use (one per use_init value) and
 * launches it; use_graph = 0 launches the same kernels directly. */
typedef struct mbrec_batch_s *mbrec_batch;
int mbrec_batch_create(mbrec_batch *batch, int ngroups, const mbrec_plan *plans, const int *group_B,
                       const float *const *data, float *const *rec, const float *const *init,
                       const float *const *init_dual, const int32_t *windows, const double *lambdaL2,
                       int num_iter);
int mbrec_batch_set_lambda(mbrec_batch batch, const double *lambdaL2);
int mbrec_batch_run(mbrec_batch batch, int use_init, int use_graph, mbrec_stream s);
int mbrec_batch_destroy(mbrec_batch batch);

/* ---- learned warm start: 4-level convolutional U-Net ----------------------
 * No counterpart in the reference: its NNLS loops start from zero
 * (nnls_L2_less.m:28-29); the network's estimate enters through the initial
 * image of mbrec_rec_nn_with_L2_reg_dev (hook at rec_nn_with_L2_reg.m:41).
 * Images are FP32 [B][H][W] device arrays, H and W multiples of 8.  Filters
 * and activations are bfloat16 on the matrix cores with FP32 accumulation;
 * master weights and Adam moments are FP32.  Parameters are one FP32 vector
 * (mbrec_unet_layout gives each layer's filter [Cout][k][k][Cin] and bias
 * offsets).  Calls on one network must not overlap. */
typedef struct mbrec_unet_s *mbrec_unet;
int mbrec_unet_create(mbrec_unet *net, int base_channels, unsigned long long seed);
int mbrec_unet_destroy(mbrec_unet net);
size_t mbrec_unet_num_params(mbrec_unet net);
int mbrec_unet_layout(mbrec_unet net, int layer, int *cin, int *cout, int *ksize, size_t *w_off, size_t *b_off);
int mbrec_unet_get_params(mbrec_unet net, float *host, size_t n);
int mbrec_unet_set_params(mbrec_unet net, const float *host, size_t n);   /* also resets Adam */
/* out = s * net(in / s), s = rms of each input image. */
int mbrec_unet_forward(mbrec_unet net, const float *in, float *out, int B, int H, int W, mbrec_stream s);
/* loss = mean((net(in/s) - target/s)^2); gradients of it (host FP32, may be NULL). */
int mbrec_unet_gradients(mbrec_unet net, const float *in, const float *target, int B, int H, int W,
                         float *grad_host, double *loss_host, mbrec_stream s);
/* on = 1: the loss (of mbrec_unet_gradients and mbrec_unet_train_step) weights item b by
 * w_b = mean_b' P_b' / P_b, P_b the mean of (target_b / s_b)^2, so that every item counts by
 * its relative error; on = 0 (the default): the plain mean.  No counterpart in the reference
 * (the warm start is new). */
int mbrec_unet_set_relative_loss(mbrec_unet net, int on);
/* One Adam step on that loss; loss_host (may be NULL) receives the loss before the step. */
int mbrec_unet_train_step(mbrec_unet net, const float *in, const float *target, int B, int H, int W, float lr,
                          float beta1, float beta2, float eps, double *loss_host, mbrec_stream s);
/* One 3x3 zero-padded convolution on the matrix cores: x bf16 [N][H][W][Cin],
 * w bf16 [Cout][3][3][Cin], y bf16 [N][H][W][Cout]; relu=1: y = relu(conv + bias),
 * relu=0: y = conv (bias must be NULL).  Cin, Cout multiples of 32. */
int mbrec_conv3x3_bf16(const uint16_t *x, const uint16_t *w, const float *bias, uint16_t *y, int N, int H, int W,
                       int Cin, int Cout, int relu, mbrec_stream s);

#ifdef __cplusplus
}
#endif
#endif /* MBREC_H */

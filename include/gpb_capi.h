/*
 * gpb_capi.h — C ABI of the MI355X gPb library (libgpb.so).
 *
 * The library computes the gPb contour signal that feeds OWT-UCM: Schmid
 * texton assignment, half-disc chi-squared cues, mPb and its oriented nonmax,
 * the intervening-contour affinity W, the generalised eigensolve
 * (D - W) v = lambda D v, sPb and gPb.  All functions return GPB_OK (0) or a
 * negative status; gpb_ctx_last_error() gives the message of the last failure
 * on a context.  A context owns one HIP device, one stream and grow-only
 * workspace; calls on one context must not overlap (use one context per host
 * thread).  Images are row-major, width = columns, height = rows.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the yanyp/UCM-MPI-GPU tree).
 */
#ifndef GPB_CAPI_H
#define GPB_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPB_OK 0
#define GPB_ERR_ARG (-1)     /* invalid argument (sizes, null pointers) */
#define GPB_ERR_HIP (-2)     /* HIP runtime error (allocation, launch) */
#define GPB_ERR_NOCONV (-3)  /* eigensolver stopped before the residual target */
#define GPB_ERR_INTERNAL (-4)
#define GPB_ERR_TOO_LARGE (-5) /* image beyond GPB_NOALLOC_MAX_PIXELS (gpb_nonmax_noalloc) */

typedef struct gpb_ctx gpb_ctx;

/* Timings and solver statistics of the last call on a context. */
typedef struct gpb_stats {
  double stage_ms[6];   /* texton, local cues, nonmax, affinity, eigensolve, sPb+gPb */
  int lanczos_matvecs;  /* operator applications (block columns for block Lanczos) */
  int lanczos_restarts;
  int lanczos_basis;
  int lanczos_block;
  int converged;        /* eigenpairs meeting the residual target */
  double max_residual;  /* max ||(D-W)v - lambda D v|| / ||D v|| over returned pairs */
  double evals[32];     /* generalised eigenvalues, ascending (first nvec valid) */
  int fallback;         /* 1 if a warm start was rejected and plain Lanczos ran */
  int reorth_second_passes; /* Gram-Schmidt steps that needed a second pass */
  int warm_mode;        /* eigensolve start: 0 plain, 1 learned block, 2 random block */
  double warm_ms;       /* warm-start network inference time (mode 1) */
  int reorth_redone_steps; /* optimistic-cycle steps repeated with the gated pass (counted in lanczos_matvecs) */
} gpb_stats;

/* Context lifetime.  Replaces the per-call device selection and
 * cudaDeviceReset of computeGPb (damascene/src/damascene.cu:258-268, :660). */
int gpb_ctx_create(int device, gpb_ctx** out);
void gpb_ctx_destroy(gpb_ctx* ctx);
const char* gpb_ctx_last_error(const gpb_ctx* ctx);

/* Number of visible GPUs.  Replaces getCudaDeviceCount
 * (damascene/include/globalPb.h:44, damascene/src/damascene.cu:242). */
int gpb_device_count(void);

/* ---- inputs -------------------------------------------------------------- */

/* Pack H*W*3 RGB bytes into 0x00BBGGRR words, the pixel format computeGPb
 * reads (src/globalPb.cpp:264-276; decode damascene/src/convert/convert.cu:19-22). */
void gpb_pack_rgb(const uint8_t* rgb, size_t npixels, uint32_t* packed);

/* Seeded Voronoi mosaic (regions, oriented sinusoid textures, Gaussian
 * noise), written as H*W*3 RGB bytes; region ids optional (may be NULL).
 * No reference counterpart: the reference reads photographs (src/main.cpp:119). */
int gpb_synth_voronoi(int width, int height, int regions, double noise_sigma, uint64_t seed, int classes,
                      uint8_t* rgb, int32_t* region);

/* 13x32 texton codebook (row-major, filter x codeword) by seeded k-means++
 * and Lloyd iterations on the image's own Schmid responses.  Replaces the
 * codebook file read by SPFeatures::genTextonMap (src/SPFeatures.cpp:317-323). */
int gpb_make_codebook(const uint32_t* packed, int width, int height, uint64_t seed, int iters, int max_samples,
                      float* codebook);

/* ---- stage (a): textons -------------------------------------------------- */

/* Texton label per pixel (0..31): Schmid 13x49x49 correlation with REFLECT
 * border and nearest codebook column.  Replaces SPFeatures::genTextonMap
 * (src/SPFeatures.cpp:277-360).  Host pointers. */
int gpb_texton(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook, uint8_t* labels);

/* Texton histogram features of superpixels, for the superpixel classifier
 * (GLSVM): the texton map of stage (a) (gpb_texton on packed, codebook), then
 * per superpixel l and texton t: descr[l][t] = 10 * (pixels of l with texton
 * t) / (pixels of l), FP64, divided then scaled as the reference does;
 * superpixels[N] holds labels in [0, nlabels) (GPB_ERR_ARG otherwise), and a
 * label with no pixels gets zeros.  Replaces SPFeatures::genTextonFeatures
 * (src/SPFeatures.cpp:185-276), whose pixel-to-codeword assignment never
 * updates its running minimum (:234-242) -- here the nearest codeword, as in
 * gpb_texton.  Host pointers. */
int gpb_texton_features(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook,
                        const int32_t* superpixels, int nlabels, double* descr);

/* Texton histograms of the window of up to 65 x 65 pixels around each centre
 * (row, col) in centres[2 * ncentres], clipped at the image border (north =
 * min(32, row), south = min(32, H - 1 - row), likewise west / east):
 * descr[k][t] = 10 * (window pixels with texton t) / (window pixels).
 * Replaces the texton part of SPFeatures::genRecPatchFeatures
 * (src/SPFeatures.cpp:478-545, textonDescrs); centres outside the image are
 * rejected with GPB_ERR_ARG.  Host pointers. */
int gpb_texton_patch_features(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook,
                              int ncentres, const int32_t* centres, double* descr);

/* ---- stage (b): local cues and mPb ------------------------------------------ */

/* Gamma 2.5, RGB->Lab (D65), normalise, quantise to 25 bins; bins[3][N]
 * (L, a, b planes) and nbins[3] = max(bin)+1 per plane.  Replaces
 * rgb_gamma_correct / rgb_to_lab / lab_normalize / quantize_values
 * (globalPb/src/math/libraries/lib_image.cc:123-245, 2315-2344) as called at
 * src/mex_pb_parts_final_selected.cpp:115-134. */
int gpb_lab_quantize(gpb_ctx* ctx, int width, int height, const uint32_t* packed, uint8_t* bins, int* nbins);

/* The 12 half-disc chi-squared cues x 8 orientations and their mPb weighted
 * sum.  cues[12][8][N] (bg r3,5,10; cga, cgb, tg r5,10,20; may be NULL) and
 * mpb_all[8][N].  Replaces det_mPb (src/det_mPb.cpp:9-22) ->
 * mex_pb_parts_final_selected (src/mex_pb_parts_final_selected.cpp:76-227) ->
 * lib_image::hist_gradient_2D (lib_image.cc:3575-3686), and the weighted sum
 * of multiscalePb (src/multiscalePb.cpp:216-245). */
int gpb_local_cues(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const uint8_t* textons, float* cues,
                   float* mpb_all);

/* Max over orientations, oriented nonmax suppression (tolerance pi/8) and the
 * clip to [0,1]: mpb_all[8][N] -> nmax[N].  Replaces nonmax_channels
 * (src/nonmax_channels.cpp:8-80), nonmax_oriented_2D (lib_image.cc:4083-4200)
 * and the clip of src/multiscalePb.cpp:251-261. */
int gpb_nonmax(gpb_ctx* ctx, int width, int height, const float* mpb_all, float* nmax);

/* gpb_nonmax on device buffers without allocating: d_mpb_all[8][N] ->
 * d_nmax[N] on the current HIP device, enqueued on stream (a hipStream_t; NULL
 * = the null stream); returns once enqueued, the caller synchronises.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc / hipFree and
 * needs no context; its scratch (5 bytes per pixel) is a file-scope device
 * array of 20 MiB that each GPU holds until the process ends, so
 * width * height <= GPB_NOALLOC_MAX_PIXELS (2048 x 2048) and height <= 65535;
 * larger images return GPB_ERR_TOO_LARGE (GPB_ERR_ARG for bad arguments,
 * GPB_ERR_HIP for a failed launch).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (two streams, two host threads) race,
 * so serialise them.  Results equal gpb_nonmax's bit for bit. */
#define GPB_NOALLOC_MAX_PIXELS (2048 * 2048)
int gpb_nonmax_noalloc(int width, int height, const float* d_mpb_all, float* d_nmax, void* stream);

/* ---- stages (c)-(e): affinity, eigenvectors, sPb, gPb ---------------------- */

/* Eigensolver settings for later calls on ctx: nvec eigenpairs (2..32,
 * default 9 as in damascene/src/damascene.cu:307), Lanczos basis size and
 * vectors kept at a restart (defaults 48 and 16), relative generalised residual target
 * ||(D-W)v - lambda D v|| / ||D v|| (default 1e-3) and a cap on operator
 * applications.  Replaces the hard-coded nEigNum/fEigTolerance of computeGPb
 * (damascene.cu:307-309) and nvec of spectralPb (src/globalPb.cpp:65). */
int gpb_set_eig_options(gpb_ctx* ctx, int nvec, int basis, int keep, double tol, int max_matvecs);

/* Convergence rule for later solves on ctx.  0 (default): a pair has
 * converged when its residual ||(D-W)v - lambda D v|| / ||D v|| is at most the
 * tol of gpb_set_eig_options (damascene's fixed tolerance, damascene.cu:309,
 * as a relative residual).  1: when that residual is at most tol times the
 * pair's Ritz value (the trivial pair: times the smallest nontrivial one), the
 * rule of the reference's CPU eigensolver (Spectra num_converged,
 * Spectra/SymEigsSolver.h:325-334, called with tol 1e-10 from
 * src/spectralPb.cpp:236).  The eigenvalues then come out of an FP64
 * Rayleigh-Ritz pass over the returned vectors. */
int gpb_set_eig_convergence(gpb_ctx* ctx, int rel);

/* Lanczos block size for later calls on ctx: 1 = single-vector thick-restart
 * Lanczos (default); 4, 8, 12 or 16 = block thick-restart Lanczos, which
 * applies W to the whole block with one read of the stencil.  No counterpart
 * in the reference (its solvers are single-vector, lanczos.cu:192-282,
 * SymEigsSolver.h:200-280). */
int gpb_set_block_size(gpb_ctx* ctx, int block);

/* ---- learned warm start (no counterpart in the reference) ------------------
 * The reference's solvers start from a fixed or random vector
 * (spectralPb.cpp:221-244, damascene/src/noReorthog/lanczos.cu:393-752).
 * Here a convolutional encoder-decoder (12 3x3 convolutions, 32..128
 * channels, BF16 on MFMA with FP32 accumulation) maps the cue features of
 * 4x4 pixel cells to 8 vectors per cell; upsampled, they start block Lanczos
 * (8 columns plus the locked D^1/2 1).  Plain Lanczos is the fallback and the
 * judge: it runs whenever the block solve fails or a returned pair's measured
 * residual exceeds the target. */

/* Eigensolve start for later gpb_compute calls: mode 0 = plain (default),
 * 1 = learned block (needs initialised parameters), 2 = block Lanczos from a
 * random block (the ablation without the network).  basis / keep: block
 * solver basis and restart sizes (<= 0: 80 / 24).  Needs nvec = 9. */
int gpb_set_warm_start(gpb_ctx* ctx, int mode, int basis, int keep);

/* Parameter vector: per layer l = 0..11, W[9][cout][cin] (tap = 3 dy + dx)
 * then bias[cout], FP32; offsets from gpb_warm_layer_info. */
size_t gpb_warm_param_count(void);
int gpb_warm_layer_info(int layer, int* cin, int* cout, int* relu, size_t* w_offset, size_t* b_offset);
/* Seeded He initialisation; resets the optimiser state. */
int gpb_warm_init(gpb_ctx* ctx, uint64_t seed);
int gpb_warm_set_params(gpb_ctx* ctx, const float* params);
int gpb_warm_get_params(gpb_ctx* ctx, float* params);

/* Training set on the device: capacity images of width x height.  _add runs
 * the path to the eigenvectors (plain Lanczos with the context's eigensolver
 * options, nvec >= 9) and stores the image's cue features and the D-weighted
 * cell sums of eigenvectors 1..8 (evals, optional: its eigenvalues). */
int gpb_warm_dataset_reset(gpb_ctx* ctx, int width, int height, int capacity);
int gpb_warm_dataset_add(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook,
                         double* evals);
int gpb_warm_dataset_size(gpb_ctx* ctx, int* n);
/* The network's cue features of an image: feat [32][ceil(H/4)][ceil(W/4)],
 * channel 8 c + o = the mean over the 4x4 cell of sum over scales s of
 * w_mpb[3 c + s] * chi2[3 c + s][o] (c: L, a, b, texton; o: orientation). */
int gpb_warm_features(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook,
                      float* feat);
/* Adam (lr, betas 0.9 / 0.999, gradient norm clipped at 1) over shuffled
 * mini-batches of training-set entries [first, first + count); augment != 0
 * puts each image through a random symmetry of the square cell grid (flips,
 * transpose; cue orientations permuted to match); epoch_loss (optional,
 * epochs entries): mean training loss per epoch. */
int gpb_warm_train(gpb_ctx* ctx, int first, int count, int epochs, int batch, float lr, uint64_t seed, int augment,
                   float* epoch_loss);
/* Mean loss (mean squared sine of the principal angles, D inner product)
 * over training-set entries [first, first + count). */
int gpb_warm_eval_loss(gpb_ctx* ctx, int first, int count, double* loss);
/* Loss and gradient for host data at cell resolution (hq, wq multiples of 8):
 * x [nimg][hq][wq][34] (32 cue features + 2 coordinates), targets
 * [nimg][hq][wq][8], dweights [nimg][hq][wq]; grad (optional, param_count),
 * y (optional, [nimg][hq][wq][8] network output). */
int gpb_warm_loss_grad(gpb_ctx* ctx, int nimg, int hq, int wq, const float* x, const float* targets,
                       const float* dweights, double* loss, float* grad, float* y);

/* SpMV tile for later solves on ctx: tile = 10 * ny + rows, a workgroup of
 * 64 x ny threads (ny wavefronts) with rows rows per thread covering 64 x
 * ny*rows pixels -- one of 42, 44, 48, 82, 84, 161, 162 -- or 0 (default) to
 * time every variant on the first solve of each image shape and keep the
 * fastest.  gpb_get_spmv_tile reports the tile in use for a shape (0 = not
 * tuned yet).  No counterpart in the reference (its SpMV has one fixed launch
 * shape, damascene/src/stencilMVM/stencilMVM.cu:129-182). */
int gpb_set_spmv_tile(gpb_ctx* ctx, int tile);
int gpb_get_spmv_tile(gpb_ctx* ctx, int width, int height, int* tile);

/* Arithmetic of the stencil SpMV for later solves on ctx: 1 (default) = the
 * differences, products and row sums in FP64, 0 = in FP32.  Operands (W, x,
 * D^-1/2) and results stay FP32 in HBM in both, so the memory traffic is the
 * same; FP64 removes the FP32 rounding that limits the relative accuracy of
 * eigenvalues near 1e-4 (smooth eigenvectors, whose row sums cancel).  The
 * reference SpMVs are FP32 (damascene/src/stencilMVM/stencilMVM.cu:129-182)
 * and FP64 (Spectra, src/spectralPb.cpp:221-244). */
int gpb_set_spmv_precision(gpb_ctx* ctx, int fp64);

/* Reorthogonalisation of the single-vector Lanczos steps for later solves on
 * ctx.  1 and 2 = partial (Simon's omega recurrence on the device: the
 * whole-basis Gram-Schmidt pass runs only when the estimated loss of
 * orthogonality exceeds sqrt(2^-24)) for residual targets >= 1e-4, and the
 * whole-basis pass every step below that; 0 = the whole-basis pass every
 * step.  With 1 the gated pass is launched on every step and exits at once
 * when not needed.  With 2 (default) a restart cycle runs without those
 * launches and is redone from the first step whose recurrence triggered,
 * with them; the results equal those of 1.  The block solver's
 * semi-orthogonal steps measure each new panel against the whole basis with
 * 1, and with 2 (default) estimate those overlaps by the block form of the
 * recurrence, measuring them only on the steps it flags; the batched solver
 * treats 1 and 2 alike.  No counterpart in the reference (Spectra
 * reorthogonalises fully, SymEigsSolver.h:200-280; damascene not at all,
 * lanczos.cu:192-282). */
int gpb_set_reorth(gpb_ctx* ctx, int mode);

/* Trigger level of the partial reorthogonalisation's omega recurrence (and
 * of the block solver's semi-orthogonality gate) for later solves on ctx
 * (0 = sqrt(2^-24), the default).  Lower levels make the gated pass run more
 * often; for verification.  No counterpart in the reference. */
int gpb_set_reorth_threshold(gpb_ctx* ctx, double level);

/* Single-vector Lanczos for later solves on ctx: 1 (default) replays each
 * restart cycle after the first from a hipGraph captured once per basis
 * buffer, 0 launches every kernel directly (the same kernels in the same
 * order).  No counterpart in the reference. */
int gpb_set_lanczos_graph(gpb_ctx* ctx, int on);

/* Single-vector Lanczos steps with partial reorthogonalisation for later
 * solves on ctx: 1 (default) fuses each three-term step into one pass over
 * the pixels -- the stencil SpMV with the previous step's scaling in its tile
 * staging and this step's update and dot products in its epilogue -- and one
 * small kernel for alpha, beta and the omega recurrence; 0 launches the
 * separate kernels (SpMV, reduction, three-term update, omega, scaling).  The
 * two differ only in FP32 rounding order.  No counterpart in the reference
 * (damascene's step is several launches with blocking host reductions,
 * lanczos.cu:192-282). */
int gpb_set_lanczos_fusion(gpb_ctx* ctx, int on);

/* Block Lanczos for later solves on ctx.  mode 0 runs each local
 * orthogonalisation pass and the dot products that follow it as separate
 * passes over the vectors; 1 fuses them (update and re-measure, update and
 * the panel's Gram, CholQR transform and the next Gram) with the same sums
 * as 0; 2 (default) is 1 with the second local Gram-Schmidt pass of a
 * semi-orthogonal step run only when a column of the panel kept less than
 * 0.01 of its norm through the first (a DGKS-style test; one pass otherwise
 * leaves local overlaps below 100 eps_32, well inside the semi-orthogonality
 * tolerance sqrt(eps_32) the next step's measuring pass enforces), and with
 * the second CholQR pass run only when the first pass' column-scaled
 * triangular factor has a Frobenius condition number of 1e3 or more.  No
 * counterpart in the reference. */
int gpb_set_block_fusion(gpb_ctx* ctx, int mode);

/* level > 0: the eigensolvers write one progress line per restart cycle to
 * stderr (operator applications, residual); 0 (default): silent. */
int gpb_set_verbose(gpb_ctx* ctx, int level);

/* Kernel classes for gpb_kernel_time. */
#define GPB_K_TEXTON 0
#define GPB_K_CUES 1
#define GPB_K_NONMAX 2
#define GPB_K_AFFINITY 3
#define GPB_K_SPMV 4
#define GPB_K_DOTS 5
#define GPB_K_UPDATE 6
#define GPB_K_RITZ 7
#define GPB_K_SPB 8
#define GPB_K_GPB 9
#define GPB_K_WARM 10
#define GPB_K_UCM 11 /* the OWT-UCM per-pixel kernels (gpb_contours2ucm_device) */

/* Enable (1) or disable (0) HIP-event timing of every kernel launch on ctx,
 * and clear the totals.  Replaces the StopWatch stage timers of computeGPb
 * (damascene/src/damascene.cu:324-575). */
int gpb_set_profiling(gpb_ctx* ctx, int on);

/* As gpb_set_profiling, for the kernel classes whose bits are set in mask
 * (bit k = class k of gpb_kernel_time; 0 disables): only those launches are
 * bracketed by events, so the others run exactly as without profiling. */
int gpb_set_profiling_classes(gpb_ctx* ctx, unsigned mask);

/* Total device time (ms, HIP events on the context stream), launch count and
 * algorithmic HBM bytes (may be NULL) of one kernel class since profiling was
 * enabled. */
int gpb_kernel_time(gpb_ctx* ctx, int kernel_class, double* total_ms, long* launches, double* algo_bytes);

/* Statistics of the last pipeline / eigensolver call. */
int gpb_get_stats(gpb_ctx* ctx, gpb_stats* out);

/* Intervening-contour affinity of the radius-5 disc (sigma 0.1), symmetrised:
 * w_upper[40][N] holds W(p, p+e_k) for the 40 offsets e_k with dy > 0 or
 * (dy == 0, dx > 0) in row-major order, 0 outside the image; degree[N] is the
 * full row sum (W_pp = 1).  mpb_nmax in [0,1].  Replaces buildW
 * (src/buildW.cpp:23-246: computeSupport, computeAffinities2, symmetrize). */
int gpb_build_w(gpb_ctx* ctx, int width, int height, const float* mpb_nmax, float* w_upper, float* degree);

/* The nvec smallest generalised eigenpairs of (D - W, D) for W built from
 * mpb_nmax: evals[nvec] ascending, evecs[nvec][N] (may be NULL) with
 * v^T D v = 1, the trivial lambda = 0 pair first.  Replaces the buildW +
 * Spectra SymGEigsSolver part of spectralPb (src/spectralPb.cpp:69-255). */
int gpb_eigs(gpb_ctx* ctx, int width, int height, const float* mpb_nmax, double* evals, float* evecs);

/* sPb[8][N] from mpb_nmax (buildW, eigensolve, min-max normalisation,
 * 1/sqrt(lambda), 8 oriented derivative-of-Gaussian filters).  Replaces
 * spectralPb (src/spectralPb.cpp:33-343); evals[nvec] may be NULL. */
int gpb_spectral_pb(gpb_ctx* ctx, int width, int height, const float* mpb_nmax, float* spb, double* evals);

/* sPb[8][N] from given eigenpairs (evecs[nvec][N], evals[nvec]); the filter
 * stage of spectralPb alone (src/spectralPb.cpp:257-322). */
int gpb_spb_from_evecs(gpb_ctx* ctx, int width, int height, int nvec, const float* evecs, const double* evals,
                       float* spb);

/* Whole gPb path on host buffers.  gpb_max[N]: max over orientations of gPb,
 * 0 where mPb_nmax <= 0.05; gpb_all[8][W][H]: calibration sigmoid of each
 * oriented gPb, column-major (index x*H + y), block i holding orientation
 * 7 - i (orientation o = the split angle o*pi/8 of the half-disc cues), the
 * block order of computeGPb's hostGPbAllConcat (oriMap {4,5,6,7,0,1,2,3},
 * damascene.cu:598-603, over filter slots s that hold orientation
 * (3 - s) mod 8, spectralPb.cu:101,706-712); gpb_orient[8][N]: the oriented
 * gPb before the sigmoid, orientation o in block o (GlobalPb's output);
 * mpb_nmax[N].  Any output may be NULL.
 * Replaces computeGPb (damascene/include/globalPb.h:37,
 * damascene/src/damascene.cu:249-661). */
int gpb_compute(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook, float* gpb_max,
                float* gpb_all, float* gpb_orient, float* mpb_nmax);

/* Same on device buffers (inputs resident in HBM, outputs device pointers,
 * any may be NULL); returns after the stream is synchronised. */
int gpb_compute_device(gpb_ctx* ctx, int width, int height, const uint32_t* d_packed, const float* d_codebook,
                       float* d_gpb_max, float* d_gpb_all, float* d_gpb_orient);

/* The spectral result of the last gpb_compute / gpb_compute_device call on
 * this context (width, height: that call's, checked): evecs[nvec][H][W]
 * (nvec <= the solve's nvec; generalised
 * eigenvectors, v^T D v = 1, eigenvalue order ascending as gpb_get_stats
 * reports them) and degree[H][W] = D = 1 + sum of the pixel's affinities;
 * either may be NULL.  Any later call that builds an affinity (eigs,
 * build_w, spectral_pb, the warm-start training set) replaces it and this
 * call then fails.  No counterpart in the reference (computeGPb keeps its
 * eigenvectors internal): it serves evaluation code that compares the
 * eigenvectors of two solves. */
int gpb_last_eigenvectors(gpb_ctx* ctx, int width, int height, int nvec, float* evecs, float* degree);

/* ---- patches: the batched form --------------------------------------------------- */

/* Cuts a width x height image into grid x grid patches (grid <= 8): interior
 * cut lines at i * size / grid, each moved by a seeded offset of up to 14% of
 * a cell in x and 10% in y, so patch sizes differ.  cores[p] = the disjoint
 * cells (x, y, w, h; they tile the image), rects[p] = each cell grown by
 * `overlap` pixels on every side that is not the image border.  Patch p is
 * row gy = p / grid, column gx = p % grid.  Replaces the equal-size split of
 * Img2Pieces (src/utilities.cpp:462-556), whose overlap is 0 (src/main.cpp:75,
 * 77). */
int gpb_cut_patches(int width, int height, int grid, int overlap, uint64_t seed, int* rects, int* cores);

/* Whole gPb of each patch rectangle of one image (rects[p] = x, y, w, h, up
 * to 64 patches), each patch an independent problem as the reference's
 * per-rank patches are (src/main.cpp:161-301).  mode 0 solves the patches'
 * eigenproblems one after another; mode 1 runs them all in one batched
 * Lanczos (the persistent batched kernels, each restart cycle replayed from a
 * hipGraph; SpMV tile and workgroup count tuned per batch shape); mode 2 is
 * mode 1 without the graph.  Outputs are stitched from each patch's core
 * rectangle cores[p] (disjoint; Pieces2Gpb, src/utilities.cpp:180-212):
 * gpb_max[H][W] (masked max over orientations) and gpb_orient[8][H][W]
 * (oriented gPb before calibration); either may be NULL.  Host buffers;
 * gpb_compute_patches_device takes device buffers (outputs left in HBM,
 * pixels outside every core untouched).  No counterpart in the reference
 * beyond the MPI patch loop it replaces. */
int gpb_compute_patches(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook, int npatch,
                        const int* rects, const int* cores, int mode, float* gpb_max, float* gpb_orient);
int gpb_compute_patches_device(gpb_ctx* ctx, int width, int height, const uint32_t* d_packed, const float* d_codebook,
                               int npatch, const int* rects, const int* cores, int mode, float* d_gpb_max,
                               float* d_gpb_orient);

/* Solver statistics and eigenvalues of patch `patch` of the last patch call;
 * gpb_get_stats then holds their sum (matvecs, converged) and maximum
 * (residual, restarts). */
int gpb_get_patch_stats(gpb_ctx* ctx, int patch, gpb_stats* out);

/* The spectral result of patch `patch` of the last patch call on ctx, for
 * patch rectangle rects[patch] (w x h): evecs[nvec][h][w] (generalised
 * eigenvectors, v' D v = 1, eigenvalue order as gpb_get_patch_stats reports
 * them), degree[h][w] (D), w_upper[40][h][w] (the affinities to the 40 upper
 * disc offsets, as gpb_build_w returns them) and mpb_nmax[h][w]; any may be
 * NULL.  A later patch call replaces them.  No counterpart in the reference
 * (each rank keeps its eigenvectors internal, src/main.cpp:161-301): it serves
 * checks of the batched solve. */
int gpb_patch_eigenvectors(gpb_ctx* ctx, int patch, int nvec, float* evecs, float* degree, float* w_upper,
                           float* mpb_nmax);

/* ---- the all-FP64 path (precision reference) ---------------------------------------- */

/* Whole gPb with every stage in FP64 on the GPU: the same texton labels, then
 * FP64 half-disc cues, mPb and nonmax, W stored in FP64 (the pair's two FP32
 * directed affinities averaged in FP64, as buildW does), an FP64 Lanczos
 * (FP64 vectors and stencil, Gram-Schmidt twice on every step) and FP64
 * sPb / gPb.  Outputs (host, any may be NULL): mpb_nmax[N], gpb_orient[8][N]
 * (before calibration), gpb_max[N] (masked max), evals[nvec], evecs[nvec][N]
 * (v' D v = 1).  The reference for the mixed-precision study of the
 * production path; the reference's own CPU path is FP64
 * (src/multiscalePb.cpp, src/spectralPb.cpp). */
int gpb_compute_f64(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook,
                    double* mpb_nmax, double* gpb_orient, double* gpb_max, double* evals, double* evecs);

/* The FP64 eigensolve alone on a given mpb_nmax (FP64 W and Lanczos):
 * evals[nvec], evecs[nvec][N] (may be NULL). */
int gpb_eigs_f64(gpb_ctx* ctx, int width, int height, const double* mpb_nmax, double* evals, double* evecs);

/* ---- OWT-UCM: the consumer of the 8 oriented gPb maps --------------------------- */

/* Ultrametric contour map from the oriented gPb maps pb_oriented[8][H][W]
 * (double, the order GlobalPb returns them in): the finest partition (the
 * maximum over orientations scaled to 8 bits, its 8-connected watershed, the
 * watershed pixels' contours subdivided into near-straight edges, each edge's
 * pixels taking the map of its chord's orientation), the doubled grid and its
 * cleaning, greedy merging of regions by mean boundary strength, and the
 * logistic calibration.  ucm: (2H+1) x (2W+1) row-major when double_size
 * (fmt "doubleSize"), else (H+1) x (W+1) (the even sites; last row and
 * column 0).  Replaces Contours2Ucm (src/contours2ucm.cpp:405-490) with
 * CreateFinestPartition (:251-404), watershed_yupeng (src/watershed.cpp:28-108),
 * fit_contour / MexContourSides (src/fit_contour.cpp:6-22,
 * src/mex_contour_sides.cpp:120-224), SuperContour4C (:24-78),
 * CleanWatersheds (:80-189), bwlabel (src/MATLAB_functions.cpp:182-222),
 * UcmMeanPb (src/ucm_mean_pb.cpp:402-449) and NormalizeOutput (:191-249).
 * Host pointers, no context (all of it is sequential host work).
 * GPB_ERR_INTERNAL where a pixel site of the cleaned grid is a boundary (the
 * reference then indexes its regions at -1). */
int gpb_contours2ucm(int width, int height, const double* pb_oriented, int double_size, double* ucm);

/* gpb_contours2ucm on oriented maps already in HBM, d_pb_oriented[8][H][W]
 * FP32 (the gpb_orient output of gpb_compute_device): the maximum over
 * orientations, its 8-bit scaling and the contour pixels' strengths run on
 * the GPU, the watershed, contour fitting, cleaning and merging on the host,
 * with the same results as gpb_contours2ucm on the same values as doubles.
 * ucm on the host.  No counterpart beyond Contours2Ucm, whose input it keeps
 * on the device. */
int gpb_contours2ucm_device(gpb_ctx* ctx, int width, int height, const float* d_pb_oriented, int double_size,
                            double* ucm);

/* The whole consumer chain of one image on ctx: gPb as gpb_compute_device
 * (device packed image and codebook), then gpb_contours2ucm_device on its
 * oriented maps; ucm on the host.  Replaces the GlobalPbCuda ->
 * UCM_generator::generateUCM2 -> Contours2Ucm sequence of one piece
 * (src/main.cpp:287, :427-458; src/UCMgenerator.cpp:23-51). */
int gpb_compute_ucm(gpb_ctx* ctx, int width, int height, const uint32_t* d_packed, const float* d_codebook,
                    int double_size, double* ucm);

/* The intermediate results of gpb_contours2ucm (each may be NULL): pb8[N]
 * (the 8-bit map), ws[N] (watershed labels, 0 = watershed pixel), ws_wt[N]
 * (finest-partition contour strength), ws_wt2 and super_ucm
 * [(2H+1)(2W+1)] (the cleaned doubled grid, the map before calibration),
 * labels[N] (initial regions), *regions, and *contours_failed (1 when the
 * contour subdivision could not be completed and the partition has no
 * contours, as the reference's exception handler leaves it), and stage_ms[7]
 * (host milliseconds: 8-bit map, watershed, contours, strength, doubled grid
 * with cleaning and regions, merging, calibration).  No counterpart beyond
 * the functions gpb_contours2ucm replaces. */
int gpb_ucm_stages(int width, int height, const double* pb_oriented, uint8_t* pb8, int32_t* ws, double* ws_wt,
                   double* ws_wt2, int32_t* labels, double* super_ucm, int* regions, int* contours_failed,
                   double* stage_ms);

/* 8-connected watershed of an 8-bit image[H][W]: levels in increasing order,
 * each level's pixels in raster order; labels[N] > 0 for basins (numbered in
 * order of creation), 0 for watershed pixels.  Replaces watershed_yupeng
 * (src/watershed.cpp:28-108). */
int gpb_watershed8(int width, int height, const uint8_t* image, int32_t* labels);

/* Contours of a binary image bw[H][W] (nonzero = contour pixel), as
 * lib_image's contour_set builds and subdivides them (globalPb/src/math/
 * libraries/lib_image.cc:5320-5485, 5953-6044, 6138-6326; called from
 * src/mex_contour_sides.cpp:136-146): vertices[2 * v] = (row, column); edges
 * [2 * e] = (start vertex, end vertex); the interior points of edge e are
 * points[2 * k] (row, column) for k in [point_offsets[e], point_offsets[e+1]).
 * Counts go to *n_vertices, *n_edges, *n_points; GPB_ERR_TOO_LARGE (with the
 * counts set) when a buffer is too small, GPB_ERR_INTERNAL when the
 * subdivision cannot be completed (the reference throws there). */
int gpb_fit_contours(int width, int height, const uint8_t* bw, int max_vertices, int32_t* vertices, int max_edges,
                     int32_t* edges, int64_t* point_offsets, int64_t max_points, int32_t* points, int* n_vertices,
                     int* n_edges, int64_t* n_points);

/* ---- matrix framing (moving outputs between processes) ------------------------- */

/* Byte framing of the reference's matrix messages for transports other than
 * MPI.  The reference sends a matrix as a header of four ints {rows, cols,
 * channels, OpenCV type} followed by rows * cols * channels elements of the
 * type's depth (MpiSendMat / MpiRecvMat, src/utilities.cpp:382-413), and a
 * vector of matrices as one int, the count, followed by those message pairs
 * (MpiSendMatVec / MpiRecvMatVec, :415-460).  A frame is those messages back to
 * back in one buffer: int32 count, then per matrix the four int32 header words
 * and the payload, host byte order, no padding (a single MpiSendMat message
 * pair is a one-matrix frame without its count).  dims[4 * i + 0..3] = rows,
 * cols, channels, type with type = depth + 8 * (channels - 1) and depth one
 * of OpenCV's 0..6 (8U, 8S, 16U, 16S, 32S, 32F, 64F; the gPb maps are CV_64FC1
 * = 6).  No counterpart beyond the MPI calls it stands in for. */
/* Bytes of the frame of count matrices (0 for invalid dims). */
size_t gpb_frame_size(int count, const int* dims);
/* Writes the frame of count matrices (payload i at data[i]) into out. */
int gpb_frame_mats(int count, const int* dims, const void* const* data, void* out, size_t out_size);
/* Reads a frame: *count matrices (GPB_ERR_ARG above max_count), their
 * dims[4 * i ..] and the byte offset of each payload in buf; GPB_ERR_ARG for a
 * truncated or malformed frame, or one with trailing bytes. */
int gpb_unframe_mats(const void* buf, size_t size, int max_count, int* count, int* dims, size_t* offsets);

/* ---- diagnostics ---------------------------------------------------------------- */

/* The intervening-contour walk program used for pixels whose radius-5 window
 * is clipped at distances (left, right, top, bottom) in [0,5] (5 = not
 * clipped): conn[81][5] per disc offset (row-major dy, dx): valid, found,
 * site count, site start (low 7 bits, high bits); sites[n][2] = (dx, dy) of
 * the pixels whose mPb maximum is the connection's intervening contour.  No
 * counterpart in the reference (it recomputes the walks per pixel,
 * buildW/src/ic.cc:375-560); exported for CPU-side cross-checks. */
int gpb_ic_program(int left, int right, int top, int bottom, int8_t* conn, int8_t* sites, int max_sites);

/* y[c] = D^-1/2 (D - W) D^-1/2 x[c] for ncols columns (x, y: [ncols][N]),
 * with W built from mpb_nmax; block = 1 applies the single-vector kernel per
 * column, block = ncols (4, 8 or 16) the block kernel.  The operator whose
 * smallest eigenpairs spectralPb needs (src/spectralPb.cpp:186-234);
 * exported for checks of the eigensolver's operator. */
int gpb_apply_laplacian(gpb_ctx* ctx, int width, int height, const float* mpb_nmax, int ncols, int block,
                        const float* x, float* y);

#ifdef __cplusplus
}
#endif

#endif /* GPB_CAPI_H */

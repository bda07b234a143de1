// Internal launch interface between the C-ABI layer and the stage kernels.
#pragma once
#include <cmath>

#include "gpb_device.hpp"

namespace gpb {
namespace owt {
struct UcmStages;
}

struct NoConvergence : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct BlockBreakdown : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Host <-> device copies of caller buffers (gpb_run.hip): through two pinned
// staging chunks, the DMA of one overlapping the host-side copy of the other
// (parallel_copy, hostcopy.cpp); small copies go directly.  stage_d2h
// returns with the data in host; stage_h2d returns once host may be reused.
void parallel_copy(void* dst, const void* src, size_t bytes);
void stage_d2h(Context& c, void* host, const void* dev, size_t bytes);
void stage_h2d(Context& c, void* dev, const void* host, size_t bytes);

// stage (a)
void upload_schmid(Context& c);
void launch_texton(Context& c, const uint32_t* d_rgb, int W, int H, const float* d_codebook, uint8_t* d_labels);

// texton histogram features (superpixel classifier)
void launch_texton_features(Context& c, const uint8_t* d_tex, int W, int H, const int32_t* d_sp, int nlabels,
                            double* d_descr);
void launch_texton_patch_features(Context& c, const uint8_t* d_tex, int W, int H, const int32_t* d_centres,
                                  int ncentres, double* d_descr);

// stage (b)
void launch_lab_quantize(Context& c, const uint32_t* d_rgb, size_t N, uint8_t* d_q /* 3N */, int* d_nb /* 3 */);
void launch_local_cues(Context& c, const uint32_t* d_rgb, const uint8_t* d_text, int W, int H, float* d_mpb_all,
                       float* d_gcue, float* d_cues /* optional [12][8][N] */,
                       float* d_feat = nullptr /* optional [32][ceil(H/4)][ceil(W/4)] cell-mean cue features */);
void launch_nonmax(Context& c, const float* d_mpb_all, int W, int H, float* d_nmax);

// stage (c): radius-5 IC stencil in a zero-padded layout (5-pixel halo)
struct Stencil {
  int W = 0, H = 0, Wp = 0, Hp = 0;
  size_t Np = 0;
  float* wup = nullptr;   // [40][Np] symmetric affinities W(p, p + e_k)
  float* D = nullptr;     // [Np] degree
  float* s = nullptr;     // [Np] D^-1/2
  int* offs = nullptr;    // [40] padded linear offsets of the upper stencil
  int off_host[kStencilHalf] = {};
  size_t pad_index(int x, int y) const { return size_t(y + kICRadius) * Wp + (x + kICRadius); }
};
// Workspace buffers are keyed by tag (distinct tags keep several stencils alive).
void build_stencil(Context& c, const float* d_mpb, int W, int H, Stencil& st, const std::string& tag = "");
void stencil_upper_offsets(int* dx, int* dy);
// The 40 upper offsets of the radius-5 disc (dy > 0, or dy == 0 and dx > 0),
// in the order the W planes are stored (tables.cpp upper_offsets; the host
// checks the two agree, spmv_offsets_match).  As compile-time constants they
// fold into the LDS read offsets of the SpMV.
struct UpperOffsets {
  int dx[kStencilHalf], dy[kStencilHalf];
};
constexpr UpperOffsets make_upper_offsets() {
  UpperOffsets t{};
  int n = 0;
  for (int u = 0; u <= kICRadius; u++)
    for (int v = -kICRadius; v <= kICRadius; v++)
      if (u * u + v * v <= kICRadius * kICRadius && (u > 0 || v > 0)) {
        t.dx[n] = v;
        t.dy[n] = u;
        n++;
      }
  return t;
}
constexpr UpperOffsets kUpper = make_upper_offsets();
bool spmv_offsets_match();
// The all-FP64 variant (precision reference): W, D, D^-1/2 stored in FP64.
struct Stencil64 {
  int W = 0, H = 0, Wp = 0, Hp = 0;
  size_t Np = 0;
  double* wup = nullptr;
  double* D = nullptr;
  double* s = nullptr;
  int* offs = nullptr;
  int off_host[kStencilHalf] = {};
};
void build_stencil_f64(Context& c, const double* d_mpb, int W, int H, Stencil64& st);
void launch_fill_f64(Context& c, double* p, size_t n, double v);
void launch_local_cues_f64(Context& c, const uint32_t* d_rgb, const uint8_t* d_text, int W, int H, double* d_mpb_all,
                           double* d_gcue, double* d_cues);
void launch_nonmax_f64(Context& c, const double* d_mpb_all, int W, int H, double* d_nmax);
// The FP32 nonmax of launch_nonmax on caller scratch (a: N floats, k: N
// bytes) and a caller stream: no context, no allocation.
void launch_nonmax_scratch(hipStream_t stream, const float* d_mpb_all, int W, int H, float* a, uint8_t* k,
                           float* d_nmax);
// Looks up this device's copy of the no-allocation scratch array (the device
// must be current); GPB_OK or GPB_ERR_HIP.
int noalloc_scratch_setup(int device);
void launch_spb_f64(Context& c, const Stencil64& st, const double* d_evecs, const double* evals, int K,
                    double* d_spb);
void launch_gpb_f64(Context& c, int W, int H, const double* d_gcue, const double* d_spb, const double* d_nmax,
                    double* d_orient, double* d_gmax);
struct LanczosStats;
void lanczos_solve_f64(Context& c, const Stencil64& st, const struct EigOptions& opt, double* evals, double* d_evecs,
                       LanczosStats& stats);
void launch_fill(Context& c, float* p, size_t n, float v);

// stage (d)
struct EigOptions {
  int K = 9;             // eigenpairs (including the trivial one)
  int m = 48;            // basis size
  int keep = 24;         // vectors kept at a restart
  double tol = 1e-3;     // relative generalised residual target
  int max_matvecs = 20000;
  // false: pair i has converged when ||(D-W)v - theta D v|| / ||D v|| <= tol;
  // true: when that residual is <= tol * max(|theta_i|, |theta_1|), the rule
  // of Spectra's num_converged (Spectra/SymEigsSolver.h:325-334: residual
  // below tol times the Ritz value), with the trivial pair held to the
  // smallest nontrivial Ritz value instead of Spectra's eps^(2/3) floor.
  bool rel = false;
};
// Residual target of pair i given the current Ritz values theta (ascending,
// K of them).
inline double conv_target(const EigOptions& o, const double* theta, int i) {
  if (!o.rel) return o.tol;
  const double t1 = std::fabs(theta[o.K > 1 ? 1 : 0]);
  const double ti = std::fabs(theta[i]);
  return o.tol * (ti > t1 ? ti : t1);
}
// Number of the K pairs whose residual meets its target.
inline int conv_count(const EigOptions& o, const double* theta, const double* resid) {
  int n = 0;
  for (int i = 0; i < o.K; i++) n += resid[i] <= conv_target(o, theta, i);
  return n;
}
// Generalised eigenpairs of (D - W, D): evals ascending (host), evecs [K][Np]
// padded layout, generalised normalisation v^T D v = 1.  start_evecs
// (optional, [start_count][Np] generalised vectors, padded, may alias d_evecs)
// replaces the random starting vector by D^1/2 times the sum of vectors
// 1 .. start_count-1: approximate eigenvectors from another solve, whose span
// the first restart cycle then recovers.
void lanczos_solve(Context& c, const Stencil& st, const EigOptions& opt, double* evals, float* d_evecs,
                   LanczosStats& stats, const float* start_evecs = nullptr, int start_count = 0);
// block form: nvec > 1 columns per Lanczos step (block thick-restart Lanczos);
// init_block (optional, [block][Np] padded, generalised or normalised-Laplacian
// vectors in u-space) replaces the random starting block.
struct BlockOptions {
  int block = 8;
  const float* init_block = nullptr;
};
void block_lanczos_solve(Context& c, const Stencil& st, const EigOptions& opt, const BlockOptions& bo, double* evals,
                         float* d_evecs, LanczosStats& stats);
void debug_block_spmv(Context& c, const Stencil& st, int B, const float* X, float* Y);
// FP64 Rayleigh-Ritz on the span of K returned pairs (generalised vectors,
// padded): refined eigenvalues (ascending) and rotated vectors, in place.
void refine_pairs(Context& c, const Stencil& st, double* evals, float* d_evecs, int K);
// Residual targets below this get the FP64 refinement: there the FP32
// storage of the Krylov vectors, not the target, limits the eigenvalues.
constexpr double kRefineBelowTol = 1e-4;
// Targets relative to the Ritz value ask for residuals of a small multiple of
// the smallest eigenvalues, always in the range where the FP32 storage limits
// the eigenvalues: they always get the refinement.  The single-vector,
// block and batched solvers keep their partial reorthogonalisation there
// (semi-orthogonality bounds the error of the projected matrix, and the
// refinement recomputes the eigenvalues from the vectors in FP64).
inline bool eig_refine(const EigOptions& o) { return o.rel || o.tol < kRefineBelowTol; }
inline bool eig_partial_reorth(const EigOptions& o) { return o.rel || o.tol >= kRefineBelowTol; }
// Warm-started solves with a target below this run the block phase to this
// residual and finish with the single-vector solver started from the block's
// Ritz vectors: block Lanczos with a few block steps per restart converges
// slowly once the wanted pairs are resolved, single-vector Lanczos does not.
constexpr double kWarmHandoverTol = 1e-3;
// Measured-residual judge of a warm-started solve whose target is relative to
// the Ritz value: plain Lanczos takes over above this.
constexpr double kWarmJudgeRel = 1e-3;
// Measured relative generalised residuals of K <= 9 returned pairs; returns
// the largest (resid, optional, receives all K).
double measured_max_residual(Context& c, const Stencil& st, const float* d_evecs, const double* evals, int K,
                             double* resid = nullptr);
// helpers shared by the solvers
void launch_reduce(Context& c, const double* partial, int G, int n, double* dst);
// launch_reduce only when *gate != 0 (dst untouched otherwise)
void launch_reduce_gated(Context& c, const double* partial, int G, int n, double* dst, const int* gate);
void launch_ritz(Context& c, const float* V, size_t ld, int m, int nk, const double* Ydev, float* out, size_t Np);
void launch_locked_start(Context& c, const Stencil& st, float* v0, float* v1_block, int B, size_t ld, uint64_t seed);
void launch_to_general(Context& c, const float* u, const float* s, float* v, size_t Np);
void launch_scale_by(Context& c, float* v, const double* nrm2, size_t Np);
// y = D^-1/2 (D - W) D^-1/2 x on padded vectors
void launch_spmv(Context& c, const Stencil& st, const float* x, float* y);
// SpMV tile code (10 * wavefronts + rows per thread) for this shape, timed
// over every variant the first time the shape is seen.
int tune_spmv(Context& c, const Stencil& st);
bool spmv_tile_valid(int code);

// stage (e)
void launch_spb(Context& c, const Stencil& st, const float* d_evecs, const double* evals, int K, float* d_spb /* [8][N] */);
void launch_gpb(Context& c, int W, int H, const float* d_gcue, const float* d_spb, const float* d_nmax,
                float* d_gpb_orient /* [8][N] or null */, float* d_gmax /* [N] or null */, float* d_gall /* [8][W][H] or null */);

// OWT-UCM of oriented gPb maps in HBM (d_pbo[8][height][width]): the
// per-pixel stages on the GPU, the sequential ones on the host
// (k_ucm.hip); ucm on the host.
void ucm_contours2ucm_device(Context& c, const float* d_pbo, int width, int height, bool double_size, double* ucm,
                             owt::UcmStages* stages);

}  // namespace gpb

// Stage (d): generalised eigensolve (D - W) v = lambda D v on the radius-5
// stencil, by thick-restart Lanczos on M = D^-1/2 W D^-1/2.
//
// Reference: spectralPb.cpp:186-244 (Spectra SymGEigsSolver, Cholesky mode,
// SymGEigsSolver.h:46-111, SymEigsSolver.h:200-586) and the GPU solver
// generalizedEigensolve / lanczos (damascene/src/noReorthog/lanczos.cu:393-752,
// 1100-), which runs up to 6000 steps without reorthogonalisation and filters
// spurious Ritz values on the host every 500 steps.
//
// Here: Lanczos runs on the normalised Laplacian L = D^-1/2 (D - W) D^-1/2,
// whose smallest eigenvalues are the wanted lambda.  The exact eigenvector
// D^1/2 1 (lambda = 0) is locked as basis vector 0;
// each step applies the FP32 stencil (SpMV), orthogonalises against the whole
// basis by classical Gram-Schmidt run twice with FP64 accumulation (dots,
// updates, norms), and keeps every scalar on the device, so a restart cycle of
// m steps has no host synchronisation.  At a restart the FP64 projected
// matrix comes to the host (Householder + QL), Ritz vectors are recombined on
// the device in FP64, and convergence is judged by the relative generalised
// residual ||(D - W) v - lambda D v|| / ||D v||, computed exactly from the
// Krylov relation (the residual of every Ritz vector is parallel to v_m).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "eig_common.hpp"

namespace gpb {


__global__ void fill_int1_kernel(int* p, int v) { *p = v; }

__global__ void fill_f32_kernel(float* __restrict__ p, size_t n, float v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

void launch_fill(Context& c, float* p, size_t n, float v) {
  if (!n) return;
  hipLaunchKernelGGL(fill_f32_kernel, dim3(std::min<unsigned>(4096, div_up(n, 256))), dim3(256), 0, c.stream, p, n, v);
  GPB_LAUNCH_CHECK();
}

template <int ROWS, typename AT, int NY>
__global__ __launch_bounds__(kSpmvTX * NY) void spmv_kernel(const float* __restrict__ wup, const float* __restrict__ s,
                                                            const float* __restrict__ x, float* __restrict__ y, int W,
                                                            int H, int Wp, int Hp, size_t Np, SpmvArgs A,
                                                            double* __restrict__ xy_partial) {
  constexpr int kSpmvTY = NY * ROWS, kSpmvTileH = kSpmvTY + 2 * kICRadius;
  __shared__ AT u[kSpmvTileH * kSpmvTileW];
  __shared__ double red[NY];
  // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin, so
  // workgroup b runs on XCD b % 8; each XCD takes a contiguous run of tiles
  // (row-major), so horizontally and vertically adjacent tiles, which read
  // each other's W rows and halo, share that XCD's L2.  (Band-major orders,
  // narrow vertical bands per XCD, measured 5-15 % slower at 2048^2 and 4096^2.)
  const int tiles_x = (W + kSpmvTX - 1) / kSpmvTX, ntiles = tiles_x * ((H + kSpmvTY - 1) / kSpmvTY);
  const int per_xcd = (ntiles + kNumXcd - 1) / kNumXcd;
  const int tile = int(blockIdx.x % kNumXcd) * per_xcd + int(blockIdx.x / kNumXcd);
  if (tile >= ntiles) return;  // whole workgroup
  const int x0 = (tile % tiles_x) * kSpmvTX, y0 = (tile / tiles_x) * kSpmvTY;  // interior coordinates of the tile origin
  const double xy = spmv_tile<ROWS, AT, NY>(wup, s, x, y, W, H, Wp, Hp, Np, A, x0, y0, u);
  if (!xy_partial) return;  // uniform
  const double t = block_sum_waves<NY>(xy, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) xy_partial[tile] = t;
}

// One fused Lanczos step over the whole image (lanczos_step_tile per tile, the
// SpMV's XCD-aware tile order): partial[k * ntiles + tile] for k = x . w,
// w . w, x . x.  a = aeff[j - 1], 1/beta and beta from beta[] at j - 1.
// For the 64x32 tile of 8 wavefronts, a register budget of 7 waves per SIMD
// (70 VGPRs, against the 80 the compiler takes unconstrained, which leaves 6)
// gives the weight stream more loads in flight: 180 -> 178 us per step at
// 2048^2.  The other shapes keep the compiler's choice.
template <int ROWS, typename AT, int NY, int XMODE>
__global__ __launch_bounds__(kSpmvTX * NY) __attribute__((amdgpu_waves_per_eu(NY == 8 && ROWS == 4 ? 7 : 1)))
void lanczos_step_kernel(
    const float* __restrict__ wup, const float* __restrict__ s, const float* __restrict__ w_in,
    const float* __restrict__ v_prev, float* __restrict__ v_out, float* __restrict__ w_out,
    const double* __restrict__ aeff, const double* __restrict__ beta, int j, int W, int H, int Wp, int Hp, size_t Np,
    double* __restrict__ partial) {
  constexpr int kSpmvTY = NY * ROWS, kSpmvTileH = kSpmvTY + 2 * kICRadius;
  __shared__ AT u[kSpmvTileH * kSpmvTileW];
  __shared__ float xs[kSpmvTileH * kSpmvTileW];
  __shared__ double red[3][NY];
  const int tiles_x = (W + kSpmvTX - 1) / kSpmvTX, ntiles = tiles_x * ((H + kSpmvTY - 1) / kSpmvTY);
  const int per_xcd = (ntiles + kNumXcd - 1) / kNumXcd;
  const int tile = int(blockIdx.x % kNumXcd) * per_xcd + int(blockIdx.x / kNumXcd);
  if (tile >= ntiles) return;  // whole workgroup
  const int x0 = (tile % tiles_x) * kSpmvTX, y0 = (tile / tiles_x) * kSpmvTY;
  const double a = XMODE == 1 ? aeff[j - 1] : 0.0, ib = XMODE == 1 ? beta[kMaxBasis + j - 1] : 0.0;
  const double b = beta[j - 1];
  double acc[3];
  lanczos_step_tile<ROWS, AT, NY, XMODE>(wup, s, w_in, v_prev, v_out, w_out, a, ib, b, W, H, Wp, Hp, Np, x0, y0, u,
                                         xs, acc);
  const int tid = threadIdx.y * kSpmvTX + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double v = acc[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((tid & 63) == 0) red[k][tid >> 6] = v;
  }
  __syncthreads();
  if (tid < 3) {
    double t = 0;
#pragma unroll
    for (int w = 0; w < NY; w++) t += red[tid][w];
    partial[size_t(tid) * ntiles + tile] = t;
  }
}

// SpMV tile codes: 10 * NY + ROWS (NY wavefronts of 64 threads, ROWS rows per
// thread; tile 64 x NY*ROWS pixels).  Every tile is at least 8 rows tall.
constexpr int kSpmvTiles[] = {42, 44, 48, 82, 84, 161, 162};
bool spmv_tile_valid(int code) {
  for (int t : kSpmvTiles)
    if (t == code) return true;
  return false;
}
static int tile_rows(int code) { return (code / 10) * (code % 10); }
int spmv_blocks(const Stencil& st, int code) { return int(div_up(st.W, kSpmvTX) * div_up(st.H, tile_rows(code))); }
int spmv_blocks(const Stencil& st) { return spmv_blocks(st, 42); }  // upper bound over the variants (8-row tiles)

bool spmv_offsets_match() {
  int dx[kStencilHalf], dy[kStencilHalf];
  stencil_upper_offsets(dx, dy);
  for (int j = 0; j < kStencilHalf; j++)
    if (dx[j] != kUpper.dx[j] || dy[j] != kUpper.dy[j]) return false;
  return true;
}

namespace {
void spmv_dispatch(Context& c, const Stencil& st, int code, const float* x, float* y, double* xy_partial) {
  static const bool offsets_ok = spmv_offsets_match();
  if (!offsets_ok) throw std::logic_error("SpMV offset table differs from the W plane order");
  SpmvArgs A;
  int dx[kStencilHalf], dy[kStencilHalf];
  stencil_upper_offsets(dx, dy);
  for (int j = 0; j < kStencilHalf; j++) {
    A.o[j] = st.off_host[j];
    A.lo[j] = dy[j] * kSpmvTileW + dx[j];
  }
  const int ntiles = spmv_blocks(st, code);
  const dim3 grid(unsigned(div_up(ntiles, kNumXcd) * kNumXcd));
#define GPB_SPMV(NY, R)                                                                                           \
  case 10 * NY + R:                                                                                              \
    if (c.spmv_fp64)                                                                                             \
      hipLaunchKernelGGL((spmv_kernel<R, double, NY>), grid, dim3(kSpmvTX, NY), 0, c.stream, st.wup, st.s, x, y, \
                         st.W, st.H, st.Wp, st.Hp, st.Np, A, xy_partial);                                        \
    else                                                                                                         \
      hipLaunchKernelGGL((spmv_kernel<R, float, NY>), grid, dim3(kSpmvTX, NY), 0, c.stream, st.wup, st.s, x, y,  \
                         st.W, st.H, st.Wp, st.Hp, st.Np, A, xy_partial);                                        \
    break;
  switch (code) {
    GPB_SPMV(4, 2)
    GPB_SPMV(4, 4)
    GPB_SPMV(4, 8)
    GPB_SPMV(8, 2)
    GPB_SPMV(8, 4)
    GPB_SPMV(16, 1)
    GPB_SPMV(16, 2)
    default:
      throw ArgError("unsupported SpMV tile");
  }
#undef GPB_SPMV
  GPB_LAUNCH_CHECK();
}
}  // namespace

// Rows per thread of the SpMV tile for this image shape: each variant is
// timed (HIP events, 3 applications after one warm-up) on the stencil itself
// the first time a shape is seen on the context; the fastest is cached.
int tune_spmv(Context& c, const Stencil& st) {
  if (c.spmv_rows_forced) return c.spmv_rows_forced;
  const uint64_t key = ((uint64_t(st.W) << 32) | uint32_t(st.H)) ^ (c.spmv_fp64 ? (1ull << 62) : 0);
  auto it = c.spmv_rows.find(key);
  if (it != c.spmv_rows.end()) return it->second;
  float* x = c.ws<float>("tune.x", st.Np);
  float* y = c.ws<float>("tune.y", st.Np);
  launch_fill(c, x, st.Np, 1.0f);
  hipEvent_t e0, e1;
  GPB_HIP(hipEventCreate(&e0));
  GPB_HIP(hipEventCreate(&e1));
  int best = 42;
  float best_ms = 1e30f;
  for (int code : kSpmvTiles) {
    spmv_dispatch(c, st, code, x, y, nullptr);
    GPB_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < 3; r++) spmv_dispatch(c, st, code, x, y, nullptr);
    GPB_HIP(hipEventRecord(e1, c.stream));
    GPB_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GPB_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (c.verbose)
      std::fprintf(stderr, "gpb: SpMV tuning %dx%d: tile %d (64x%d): %.1f us\n", st.W, st.H, code, tile_rows(code),
                   ms * 1e3f / 3);
    if (ms < best_ms) {
      best_ms = ms;
      best = code;
    }
  }
  GPB_HIP(hipEventDestroy(e0));
  GPB_HIP(hipEventDestroy(e1));
  c.spmv_rows[key] = best;
  return best;
}

// y = L x; with xy_partial, also the per-tile partial sums of x . y
// (spmv_blocks(st, code) of them, code = the tile the shape was tuned to).
void launch_spmv(Context& c, const Stencil& st, const float* x, float* y, double* xy_partial) {
  const int code = tune_spmv(c, st);
  KScope ks(c, kKSpmv, double(st.W) * st.H * (4 * kStencilHalf + 12));
  spmv_dispatch(c, st, code, x, y, xy_partial);
}

void launch_spmv(Context& c, const Stencil& st, const float* x, float* y) { launch_spmv(c, st, x, y, nullptr); }

static void lanczos_step_dispatch(Context& c, const Stencil& st, int code, int xmode, const float* w_in,
                                  const float* v_prev, float* v_out, float* w_out, const double* aeff,
                                  const double* beta, int j, double* partial);

__global__ void fill_f64_kernel(double* p, int n, double v) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// Tile code of the fused step for this shape: every variant timed (HIP
// events, 3 steps after one warm-up) the first time a shape is seen, on
// scratch vectors; the fastest is cached (its own choice: the fused kernel's
// LDS and register needs differ from the plain SpMV's).
static int tune_lanczos_step(Context& c, const Stencil& st) {
  if (c.spmv_rows_forced) return c.spmv_rows_forced;
  const uint64_t key = ((uint64_t(st.W) << 32) | uint32_t(st.H)) ^ (c.spmv_fp64 ? (1ull << 62) : 0) ^ (1ull << 61);
  auto it = c.spmv_rows.find(key);
  if (it != c.spmv_rows.end()) return it->second;
  float* win = c.ws<float>("tune.x", st.Np);
  float* vp = c.ws<float>("tune.y", st.Np);
  float* vo = c.ws<float>("tune.v", st.Np);
  float* wo = c.ws<float>("tune.w", st.Np);
  double* bt = c.ws<double>("tune.beta", 2 * kMaxBasis);
  double* ae = c.ws<double>("tune.aeff", kMaxBasis);
  double* part = c.ws<double>("tune.part", size_t(3) * spmv_blocks(st));
  launch_fill(c, win, st.Np, 1.0f);
  launch_fill(c, vp, st.Np, 0.0f);
  hipLaunchKernelGGL(fill_f64_kernel, dim3(1), dim3(256), 0, c.stream, bt, 2 * kMaxBasis, 1.0);
  hipLaunchKernelGGL(fill_f64_kernel, dim3(1), dim3(256), 0, c.stream, ae, kMaxBasis, 0.0);
  GPB_LAUNCH_CHECK();
  hipEvent_t e0, e1;
  GPB_HIP(hipEventCreate(&e0));
  GPB_HIP(hipEventCreate(&e1));
  int best = 42;
  float best_ms = 1e30f;
  for (int code : kSpmvTiles) {
    lanczos_step_dispatch(c, st, code, 1, win, vp, vo, wo, ae, bt, 1, part);
    GPB_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < 3; r++) lanczos_step_dispatch(c, st, code, 1, win, vp, vo, wo, ae, bt, 1, part);
    GPB_HIP(hipEventRecord(e1, c.stream));
    GPB_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GPB_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (c.verbose)
      std::fprintf(stderr, "gpb: fused-step tuning %dx%d: tile %d (64x%d): %.1f us\n", st.W, st.H, code,
                   tile_rows(code), ms * 1e3f / 3);
    if (ms < best_ms) {
      best_ms = ms;
      best = code;
    }
  }
  GPB_HIP(hipEventDestroy(e0));
  GPB_HIP(hipEventDestroy(e1));
  c.spmv_rows[key] = best;
  return best;
}

// One fused step (lanczos_step_kernel) with the tile tuned for it; xmode 1
// forms v_j from w_in and v_{j-1}, 0 reads it from v_out.
static void launch_lanczos_step(Context& c, const Stencil& st, int xmode, const float* w_in, const float* v_prev,
                                float* v_out, float* w_out, const double* aeff, const double* beta, int j,
                                double* partial) {
  const int code = tune_lanczos_step(c, st);
  // W planes, w_in, v_{j-1}, s read; v_j (xmode 1) and w written
  KScope ks(c, kKSpmv, double(st.W) * st.H * (4 * kStencilHalf + (xmode ? 20 : 16)));
  lanczos_step_dispatch(c, st, code, xmode, w_in, v_prev, v_out, w_out, aeff, beta, j, partial);
}

static void lanczos_step_dispatch(Context& c, const Stencil& st, int code, int xmode, const float* w_in,
                                  const float* v_prev, float* v_out, float* w_out, const double* aeff,
                                  const double* beta, int j, double* partial) {
  const int ntiles = spmv_blocks(st, code);
  const dim3 grid(unsigned(div_up(ntiles, kNumXcd) * kNumXcd));
#define GPB_STEP(NY, R, XM)                                                                                          \
  if (c.spmv_fp64)                                                                                                 \
    hipLaunchKernelGGL((lanczos_step_kernel<R, double, NY, XM>), grid, dim3(kSpmvTX, NY), 0, c.stream, st.wup, st.s, \
                       w_in, v_prev, v_out, w_out, aeff, beta, j, st.W, st.H, st.Wp, st.Hp, st.Np, partial);       \
  else                                                                                                             \
    hipLaunchKernelGGL((lanczos_step_kernel<R, float, NY, XM>), grid, dim3(kSpmvTX, NY), 0, c.stream, st.wup, st.s,  \
                       w_in, v_prev, v_out, w_out, aeff, beta, j, st.W, st.H, st.Wp, st.Hp, st.Np, partial);
#define GPB_STEP_CODE(NY, R)  \
  case 10 * NY + R:           \
    if (xmode) {              \
      GPB_STEP(NY, R, 1)      \
    } else {                  \
      GPB_STEP(NY, R, 0)      \
    }                         \
    break;
  switch (code) {
    GPB_STEP_CODE(4, 2)
    GPB_STEP_CODE(4, 4)
    GPB_STEP_CODE(4, 8)
    GPB_STEP_CODE(8, 2)
    GPB_STEP_CODE(8, 4)
    GPB_STEP_CODE(16, 1)
    GPB_STEP_CODE(16, 2)
    default:
      throw ArgError("unsupported SpMV tile");
  }
#undef GPB_STEP_CODE
#undef GPB_STEP
  GPB_LAUNCH_CHECK();
}

// Sum of 4 values over the 256-thread block; result valid in thread 0.
__device__ __forceinline__ void block_sum4(double v[4], double* sh /* [4][4] */) {
#pragma unroll
  for (int k = 0; k < 4; k++)
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) sh[wv * 4 + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = sh[k] + sh[4 + k] + sh[8 + k] + sh[12 + k];
  __syncthreads();
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  double a[4] = {v, 0, 0, 0};
  block_sum4(a, sh);
  return a[0];
}


// Dot products as a 2-D grid: block (c, i) reduces V_i . w over element chunk
// c (kDotChunk elements, 16 per thread) and writes partial[i][c]; column n is
// w . (D) w when wsq.  Every block streams one contiguous chunk of one vector
// and reduces once, so the grid saturates HBM; w is re-read from L2.
constexpr int kDotChunk = 4096;

__global__ __launch_bounds__(256) void multidot2_kernel(const float* __restrict__ V, size_t ld, int n,
                                                        const float* __restrict__ w, const float* __restrict__ D,
                                                        size_t Np, double* __restrict__ partial) {
  __shared__ double sh[16];
  const int i = blockIdx.y;
  const size_t lo = size_t(blockIdx.x) * kDotChunk;
  const float* v = (i < n) ? V + size_t(i) * ld : w;
  const bool weighted = (i == n) && D;
  double acc = 0;
#pragma unroll 4
  for (int e = threadIdx.x; e < kDotChunk; e += 256) {
    const size_t p = lo + e;
    if (p < Np) {
      const double x = v[p], y = w[p];
      acc += (weighted ? double(D[p]) : 1.0) * x * y;
    }
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[size_t(i) * gridDim.x + blockIdx.x] = r;
}

// w <- w - sum_i h_i V_i (FP64 accumulation), grid-stride; per-block partial
// of ||w||^2 into nrm_partial[blk] when non-null.  NV >= n coefficients are
// held in registers and all NV basis loads of an element are issued before
// use (rows past n re-read row 0 with coefficient 0).  Skipped when
// *gate == 0.
template <int NV>
__global__ __launch_bounds__(256) void update_gs_kernel(const float* __restrict__ V, size_t ld, int n,
                                                        const double* __restrict__ h, float* __restrict__ w, size_t Np,
                                                        double* __restrict__ nrm_partial, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double sh[16];
  double hv[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) hv[i] = i < n ? h[i] : 0.0;
  double nrm = 0;
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    const float* q = V + p;
    float vv[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) vv[i] = q[size_t(i < n ? i : 0) * ld];
    double acc = w[p];
#pragma unroll
    for (int i = 0; i < NV; i++) acc -= hv[i] * double(vv[i]);
    const float f = float(acc);
    w[p] = f;
    nrm += double(f) * double(f);
  }
  if (nrm_partial) {
    const double r = block_sum(nrm, sh);
    if (threadIdx.x == 0) nrm_partial[blockIdx.x] = r;
  }
}

// The same update for bases wider than the register version (n > 64).
__global__ __launch_bounds__(256) void update_wide_kernel(const float* __restrict__ V, size_t ld, int n,
                                                          const double* __restrict__ h, float* __restrict__ w,
                                                          size_t Np, double* __restrict__ nrm_partial,
                                                          const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double hs[kMaxBasis];
  __shared__ double sh[16];
  for (int i = threadIdx.x; i < n; i += blockDim.x) hs[i] = h[i];
  __syncthreads();
  double nrm = 0;
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    double acc = w[p];
    for (int i = 0; i < n; i++) acc -= hs[i] * double(V[size_t(i) * ld + p]);
    const float f = float(acc);
    w[p] = f;
    nrm += double(f) * double(f);
  }
  if (nrm_partial) {
    const double r = block_sum(nrm, sh);
    if (threadIdx.x == 0) nrm_partial[blockIdx.x] = r;
  }
}

void launch_update(Context& c, const float* V, size_t ld, int n, const double* h, float* w, size_t Np,
                   double* nrm_partial, const int* gate, int G) {
  if (n < 1 || n > kProjectMaxN) throw ArgError("update width out of range");
  const dim3 g(G), b(256);
#define GPB_UPD(NV)                                                                                               \
  case NV / 4:                                                                                                    \
    hipLaunchKernelGGL(update_gs_kernel<NV>, g, b, 0, c.stream, V, ld, n, h, w, Np, nrm_partial, gate);         \
    break;
  switch ((n + 3) / 4) {
    GPB_UPD(4)
    GPB_UPD(8)
    GPB_UPD(12)
    GPB_UPD(16)
    GPB_UPD(20)
    GPB_UPD(24)
    GPB_UPD(28)
    GPB_UPD(32)
    GPB_UPD(36)
    GPB_UPD(40)
    GPB_UPD(44)
    GPB_UPD(48)
    GPB_UPD(52)
    GPB_UPD(56)
    GPB_UPD(60)
    GPB_UPD(64)
  }
#undef GPB_UPD
  GPB_LAUNCH_CHECK();
}

// Lanczos projection in one pass over the basis.  With alpha: first the
// three-term step w <- w - alpha v_j - beta_{j-1} v_{j-1} (FP64, rounded to
// FP32 and stored), then, for the result, the dots V_i . w (i < n) and
// ||w||^2, FP64 accumulation in NV per-thread registers, so w is read once
// however many basis vectors there are.  partial[i][blk] for i < n, the
// squared norm in row n.  Skipped when *gate == 0.
template <int NV>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ V, size_t ld, int n,
                                                      float* __restrict__ w, const double* __restrict__ alpha,
                                                      const double* __restrict__ beta, int j,
                                                      const int* __restrict__ gate, size_t Np,
                                                      double* __restrict__ partial) {
  if (gate && *gate == 0) return;
  __shared__ double sh[4][NV + 1];
  const double a = alpha ? *alpha : 0.0;
  const double b = (alpha && j > 0) ? beta[j - 1] : 0.0;
  const float* vj = V + size_t(j) * ld;
  const float* vjm = V + size_t(j > 0 ? j - 1 : 0) * ld;
  double acc[NV + 1];
#pragma unroll
  for (int i = 0; i <= NV; i++) acc[i] = 0.0;
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    double wn = w[p];
    if (alpha) {
      wn -= a * double(vj[p]) + b * double(vjm[p]);
      const float f = float(wn);
      w[p] = f;
      wn = f;
    }
    acc[NV] += wn * wn;
    // all NV loads issued before any use (rows past n re-read row 0 and are
    // discarded), so the wave keeps NV loads in flight
    const float* q = V + p;
    float vv[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) vv[i] = q[size_t(i < n ? i : 0) * ld];
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] += double(vv[i]) * wn;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i <= NV; i++) {
    if (i < n || i == NV) {
      double v = acc[i];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) sh[wv][i] = v;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t <= n; t += blockDim.x) {
    const int i = t < n ? t : NV;
    partial[size_t(t) * gridDim.x + blockIdx.x] = sh[0][i] + sh[1][i] + sh[2][i] + sh[3][i];
  }
}

void launch_project(Context& c, const float* V, size_t ld, int n, float* w, const double* alpha, const double* beta,
                    int j, const int* gate, size_t Np, double* partial, int G) {
  if (n < 1 || n > kProjectMaxN) throw ArgError("projection width out of range");
  const dim3 g(G), b(256);
#define GPB_PROJ(NV)                                                                                                \
  case NV / 4:                                                                                                      \
    hipLaunchKernelGGL(project_kernel<NV>, g, b, 0, c.stream, V, ld, n, w, alpha, beta, j, gate, Np, partial);      \
    break;
  switch ((n + 3) / 4) {
    GPB_PROJ(4)
    GPB_PROJ(8)
    GPB_PROJ(12)
    GPB_PROJ(16)
    GPB_PROJ(20)
    GPB_PROJ(24)
    GPB_PROJ(28)
    GPB_PROJ(32)
    GPB_PROJ(36)
    GPB_PROJ(40)
    GPB_PROJ(44)
    GPB_PROJ(48)
    GPB_PROJ(52)
    GPB_PROJ(56)
    GPB_PROJ(60)
    GPB_PROJ(64)
  }
#undef GPB_PROJ
  GPB_LAUNCH_CHECK();
}

// Classical Gram-Schmidt's update followed by the next pass's projection,
// in one pass over the basis: w <- fl32(w - sum_i h_i V_i) (the update's
// arithmetic), then the dots V_i . w and ||w||^2 of the new w with the
// projection's per-thread accumulation and block reduction, so the partials
// equal those of update_gs_kernel followed by project_kernel on the same grid
// (launch it on the projection's grid).
template <int NV>
__global__ __launch_bounds__(256) void update_project_kernel(const float* __restrict__ V, size_t ld, int n,
                                                             const double* __restrict__ h, float* __restrict__ w,
                                                             size_t Np, double* __restrict__ partial) {
  __shared__ double sh[4][NV + 1];
  double hv[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) hv[i] = i < n ? h[i] : 0.0;
  double acc[NV + 1];
#pragma unroll
  for (int i = 0; i <= NV; i++) acc[i] = 0.0;
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    const float* q = V + p;
    float vv[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) vv[i] = q[size_t(i < n ? i : 0) * ld];
    double u = w[p];
#pragma unroll
    for (int i = 0; i < NV; i++) u -= hv[i] * double(vv[i]);
    const float f = float(u);
    w[p] = f;
    const double wn = f;
    acc[NV] += wn * wn;
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] += double(vv[i]) * wn;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i <= NV; i++) {
    if (i < n || i == NV) {
      double v = acc[i];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) sh[wv][i] = v;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t <= n; t += blockDim.x) {
    const int i = t < n ? t : NV;
    partial[size_t(t) * gridDim.x + blockIdx.x] = sh[0][i] + sh[1][i] + sh[2][i] + sh[3][i];
  }
}

void launch_update_project(Context& c, const float* V, size_t ld, int n, const double* h, float* w, size_t Np,
                           double* partial, int G) {
  if (n < 1 || n > kProjectMaxN) throw ArgError("projection width out of range");
  const dim3 g(G), b(256);
#define GPB_UP(NV)                                                                                               \
  case NV / 4:                                                                                                   \
    hipLaunchKernelGGL(update_project_kernel<NV>, g, b, 0, c.stream, V, ld, n, h, w, Np, partial);               \
    break;
  switch ((n + 3) / 4) {
    GPB_UP(4)
    GPB_UP(8)
    GPB_UP(12)
    GPB_UP(16)
    GPB_UP(20)
    GPB_UP(24)
    GPB_UP(28)
    GPB_UP(32)
    GPB_UP(36)
    GPB_UP(40)
    GPB_UP(44)
    GPB_UP(48)
    GPB_UP(52)
    GPB_UP(56)
    GPB_UP(60)
    GPB_UP(64)
  }
#undef GPB_UP
  GPB_LAUNCH_CHECK();
}

// Weighted squared norms v_k . D v_k of the vectors V + sel_k ld (k < m) in
// one launch: block (chunk, k) with multidot2_kernel's chunking and
// arithmetic, so each equals that kernel's weighted norm of the same vector.
struct NormSel {
  int idx[kMaxNvec + 2];
};
__global__ __launch_bounds__(256) void wnorms_kernel(const float* __restrict__ V, size_t ld, NormSel sel,
                                                     const float* __restrict__ D, size_t Np,
                                                     double* __restrict__ partial) {
  __shared__ double sh[16];
  const int k = blockIdx.y;
  const size_t lo = size_t(blockIdx.x) * kDotChunk;
  const float* v = V + size_t(sel.idx[k]) * ld;
  double acc = 0;
#pragma unroll 4
  for (int e = threadIdx.x; e < kDotChunk; e += 256) {
    const size_t p = lo + e;
    if (p < Np) {
      const double x = v[p];
      acc += double(D[p]) * x * x;
    }
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[size_t(k) * gridDim.x + blockIdx.x] = r;
}

// out[i] = sum_blk partial[i][blk]  (one block per i, fixed order); gated.
__global__ __launch_bounds__(kRedThreads) void reduce_kernel(const double* __restrict__ partial, int G,
                                                             double* __restrict__ out, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double sh[16];
  double a;
  strided_sums<1, kRedThreads>(partial + size_t(blockIdx.x) * G, G, threadIdx.x, &a);
  const double r = block_sum(a, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}



// Daniel-Gragg-Kaufman-Stewart test (1976): after the projection pass a
// second pass runs only when the norm dropped below 1/sqrt(2) of its value
// before the pass (w1sq < w0sq / 2).  acc[0], acc[1]: algorithmic bytes of the
// gated projection / update launches that run; acc[2]: second passes.
__global__ void reorth_decide_kernel(const double* __restrict__ w0sq, const double* __restrict__ w1sq,
                                     int* __restrict__ flag, int n, double np, double* __restrict__ acc) {
  const int f = (*w1sq < 0.5 * *w0sq) ? 1 : 0;
  *flag = f;
  if (f) {
    acc[0] += np * (4.0 * n + 4);
    acc[1] += np * (4.0 * n + 8);
    acc[2] += 1.0;
  }
}

// Column j of the projected matrix: T[i][j] += h1[i] (+ h2[i] after a second
// pass) (+ alpha on the diagonal), mirrored into row j; beta_j from the norm
// after the last pass that ran, stored in beta[j] (1/beta_j in
// beta[kMaxBasis + j]) and, when couple_next, as T[j][j+1] = T[j+1][j], the
// coupling the next step's three-term recurrence uses (a next step that
// projects against the whole basis finds it itself).
__global__ void tcol_kernel(double* __restrict__ T, int m, int j, int couple_next, const double* __restrict__ alpha,
                            const double* __restrict__ h1, const double* __restrict__ h2, const int* __restrict__ flag,
                            const double* __restrict__ nrm1, const double* __restrict__ nrm2,
                            double* __restrict__ beta) {
  const int i = threadIdx.x;
  const bool two = flag && *flag;
  if (i <= j) {
    const double v = T[size_t(i) * m + j] + h1[i] + (two ? h2[i] : 0.0) + ((i == j && alpha) ? *alpha : 0.0);
    T[size_t(i) * m + j] = v;
    if (i != j) T[size_t(j) * m + i] = v;
  }
  if (i == 0 && beta) {
    const double b = sqrt(two ? *nrm2 : *nrm1);
    beta[j] = b;
    beta[kMaxBasis + j] = b > 1e-300 ? 1.0 / b : 0.0;
    if (couple_next && j + 1 < m) {
      T[size_t(j) * m + j + 1] = b;
      T[size_t(j + 1) * m + j] = b;
    }
  }
}

// ---- partial reorthogonalisation (Simon 1984) ---------------------------------
// Semi-orthogonal Lanczos: a step applies the three-term recurrence only, and
// the omega recurrence estimates omega_{j+1,k} = v_{j+1} . v_k from the
// projected matrix; the whole-basis pass runs only when some estimate exceeds
// sqrt(eps) (and once more on the step after, whose estimate depends on the
// pre-correction vector too).  eps is the unit roundoff of the FP32-stored
// basis.  T is the dense projected matrix (tridiagonal plus the arrowhead of a
// thick restart), so (T omega_j)_k also covers the kept Ritz vectors.
constexpr double kProEps = 5.9604644775390625e-08;  // 2^-24
constexpr double kProSqrtEps = 2.44140625e-04;      // 2^-12

// w <- w - alpha v_j - beta_{j-1} v_{j-1} (FP64, stored FP32); per-block
// partials of ||w||^2.
__global__ __launch_bounds__(256) void three_term_kernel(const float* __restrict__ V, size_t ld, int j,
                                                         const double* __restrict__ alpha,
                                                         const double* __restrict__ beta, float* __restrict__ w,
                                                         size_t Np, double* __restrict__ nrm_partial) {
  __shared__ double sh[16];
  const double a = *alpha, b = beta[j - 1];
  const float* vj = V + size_t(j) * ld;
  const float* vjm = V + size_t(j - 1) * ld;
  double nrm = 0;
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    const float f = float(double(w[p]) - a * double(vj[p]) - b * double(vjm[p]));
    w[p] = f;
    nrm += double(f) * double(f);
  }
  const double r = block_sum(nrm, sh);
  if (threadIdx.x == 0) nrm_partial[blockIdx.x] = r;
}

// Step j (one block of 256 threads, j < m < kMaxBasis): beta_j from the
// partial norms, the omega recurrence for k < j,
//   b_new om1[k] = (T om0)_k - alpha_j om0[k] - b_cur om_1[k] +- eps (b_new + |T[k][k+1]|),
// with om0 = omega_j (om0[j] = 1), om_1 = omega_{j-1} (om_1[j-1] = 1); the
// reorthogonalisation gate; T[j][j] = alpha and, without a reorthogonalisation,
// beta[j] and the coupling T[j][j+1].  om1 overwrites om_1 in place.
// acc: byte and pass counters as for the DGKS gate.
//
// omp holds omega_{j-1}, om0 omega_j; omega_{j+1} goes to om_1 (which may
// alias omp: each thread reads its entry before writing it).  thr is the
// trigger level (sqrt(eps) by default).  With first_viol (the optimistic
// cycle, no gated launches), the first step whose omega exceeds thr is
// recorded there and the pass counters are left alone.
// The omega recurrence and the gate for step j, given alpha_j (al) and
// beta_j (bn); one block of 256 threads.  aeff (optional): receives alpha_j
// when no whole-basis pass follows (the fused steps apply it in the next
// step's staging).
__device__ void omega_body(double* __restrict__ T, int m, int j, double al, double bn, double* __restrict__ beta,
                           const double* __restrict__ om0, const double* omp, double* om_1, int* __restrict__ flag,
                           int* __restrict__ pending, double np, double* __restrict__ acc, double thr,
                           int* __restrict__ first_viol, double* __restrict__ aeff, double* mx) {
  const int t = threadIdx.x;
  const double bc = beta[j - 1];
  double om = 0;
  if (t < j) {
    const double tw = dot_in_order(T + size_t(t) * m, om0, j, T[size_t(t) * m + j]);  // om0[j] = 1
    const double prev = (t == j - 1) ? 1.0 : omp[t];
    const double e = tw - al * om0[t] - bc * prev;
    const double noise = kProEps * (bn + fabs(T[size_t(t) * m + t + 1]));
    om = (e + (e >= 0 ? noise : -noise)) / (bn > 1e-300 ? bn : 1e-300);
  }
  double v = fabs(om);
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((t & 63) == 0) mx[t >> 6] = v;
  __syncthreads();
  const double worst = fmax(fmax(mx[0], mx[1]), fmax(mx[2], mx[3]));
  const int pend = *pending;
  const bool viol = worst > thr;
  const bool f = viol || pend;
  // after a reorthogonalisation the new vector is orthogonal to rounding level
  if (t < j) om_1[t] = f ? kProEps : om;
  if (t == j) om_1[t] = kProEps;
  if (t == 0) {
    T[size_t(j) * m + j] = al;
    *flag = f ? 1 : 0;
    *pending = (viol && !pend) ? 1 : 0;
    if (first_viol && viol && *first_viol > j) *first_viol = j;
    if (f && !first_viol) {
      acc[0] += np * (4.0 * (j + 1) + 4);
      acc[1] += np * (4.0 * (j + 1) + 8);
      acc[2] += 1.0;
    } else if (!f) {
      beta[j] = bn;
      beta[kMaxBasis + j] = bn > 1e-300 ? 1.0 / bn : 0.0;
      if (aeff) aeff[j] = al;
      if (j + 1 < m) {
        T[size_t(j) * m + j + 1] = bn;
        T[size_t(j + 1) * m + j] = bn;
      }
    }
  }
}

__global__ __launch_bounds__(256) void omega_kernel(double* __restrict__ T, int m, int j,
                                                    const double* __restrict__ alpha, double* __restrict__ beta,
                                                    const double* __restrict__ nrm_partial, int G,
                                                    const double* __restrict__ om0, const double* omp, double* om_1,
                                                    int* __restrict__ flag, int* __restrict__ pending, double np,
                                                    double* __restrict__ acc, double thr, int* __restrict__ first_viol) {
  __shared__ double red4[4];
  __shared__ double mx[4];
  double a;
  strided_sums<1, 256>(nrm_partial, G, threadIdx.x, &a);
  const double nrm = block_sum_waves<4>(a, red4);
  omega_body(T, m, j, *alpha, sqrt(nrm), beta, om0, omp, om_1, flag, pending, np, acc, thr, first_viol, nullptr, mx);
}

// Step j of the fused form: alpha_j = x . w and beta_j^2 = ||w - alpha x||^2 =
// w . w - 2 alpha (x . w) + alpha^2 (x . x) from the tile partials of the
// fused kernel (partial[3][G], summed in a fixed order), then the omega
// recurrence and the gate.  alpha_j also goes to alpha_out for a gated pass.
__global__ __launch_bounds__(256) void omega_fused_kernel(double* __restrict__ T, int m, int j,
                                                          const double* __restrict__ partial, int G,
                                                          double* __restrict__ beta, double* __restrict__ aeff,
                                                          double* __restrict__ alpha_out,
                                                          const double* __restrict__ om0, const double* omp,
                                                          double* om_1, int* __restrict__ flag,
                                                          int* __restrict__ pending, double np,
                                                          double* __restrict__ acc, double thr,
                                                          int* __restrict__ first_viol) {
  __shared__ double red4[4];
  __shared__ double mx[4];
  double s3[3];
  strided_sums<3, 256>(partial, G, threadIdx.x, s3);
  const double xw = block_sum_waves<4>(s3[0], red4);
  __syncthreads();
  const double ww = block_sum_waves<4>(s3[1], red4);
  __syncthreads();
  const double xx = block_sum_waves<4>(s3[2], red4);
  const double al = xw;
  const double b2 = ww - 2.0 * al * xw + al * al * xx;
  const double bn = sqrt(b2 > 0 ? b2 : 0.0);
  if (threadIdx.x == 0) *alpha_out = al;
  omega_body(T, m, j, al, bn, beta, om0, omp, om_1, flag, pending, np, acc, thr, first_viol, aeff, mx);
}

// Gated whole-basis pass of a fused step, first and last parts: w -= alpha v_j
// (the update the fused form leaves to the next step), and after the
// projection beta_j from the norm that pass left, with alpha_eff = 0 (the next
// step's staging then only scales).  Skipped when *flag == 0.
__global__ __launch_bounds__(256) void materialise_kernel(float* __restrict__ w, const float* __restrict__ vj,
                                                          const double* __restrict__ alpha, size_t Np,
                                                          const int* __restrict__ flag) {
  if (*flag == 0) return;
  const double a = *alpha;
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    w[p] = float(double(w[p]) - a * double(vj[p]));
}

__global__ void gated_beta_kernel(double* __restrict__ T, int m, int j, double* __restrict__ beta,
                                  double* __restrict__ aeff, const double* __restrict__ nrm2,
                                  const int* __restrict__ flag) {
  if (*flag == 0) return;
  const double bn = sqrt(*nrm2);
  beta[j] = bn;
  beta[kMaxBasis + j] = bn > 1e-300 ? 1.0 / bn : 0.0;
  aeff[j] = 0.0;
  if (j + 1 < m) {
    T[size_t(j) * m + j + 1] = bn;
    T[size_t(j + 1) * m + j] = bn;
  }
}

// The residual direction at the end of a fused cycle: v_m = (w - alpha_eff
// v_{m-1}) / beta_{m-1}, the scaling the next step would have done.
__global__ __launch_bounds__(256) void fused_last_kernel(const float* __restrict__ w, const float* __restrict__ vprev,
                                                         const double* __restrict__ aeff,
                                                         const double* __restrict__ beta, int j,
                                                         float* __restrict__ v, size_t Np) {
  const double a = aeff[j], ib = beta[kMaxBasis + j];
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    v[p] = float((double(w[p]) - a * double(vprev[p])) * ib);
}

// omega_{j+1}[k] = omega_j[k] = eps for k <= j after a full
// Gram-Schmidt step j (the first of a cycle); pending cleared.
__global__ void omega_reset_kernel(double* __restrict__ om_a, double* __restrict__ om_b, int j,
                                   int* __restrict__ pending, double* __restrict__ aeff) {
  if (threadIdx.x == 0) aeff[j] = 0.0;  // this step's w is fully orthogonalised: the next staging only scales
  for (int k = threadIdx.x; k <= j; k += blockDim.x) {
    om_a[k] = kProEps;
    om_b[k] = kProEps;
  }
  if (threadIdx.x == 0) *pending = 0;
}

__global__ __launch_bounds__(256) void scale_kernel(const float* __restrict__ w, const double* __restrict__ beta, int j,
                                                    float* __restrict__ v, size_t Np) {
  const double ib = beta[kMaxBasis + j];
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    v[p] = float(double(w[p]) * ib);
}

// Last launch of a partially reorthogonalised step: v_{j+1} = w / beta_j.
// When the gated whole-basis pass ran (*flag), every workgroup takes beta_j
// from the norm that pass left (the same expression in each) and workgroup 0
// records it in beta and T; otherwise beta_j is the one omega_kernel recorded.
// The components the gated pass removed stay out of T, which keeps the
// tridiagonal form the three-term steps and the omega recurrence assume
// (recording them at isolated steps made the restarted iteration diverge).
__global__ __launch_bounds__(256) void pro_scale_kernel(const float* __restrict__ w, double* __restrict__ beta, int j,
                                                        float* __restrict__ v, size_t Np, double* __restrict__ T,
                                                        int m, const double* __restrict__ nrm2,
                                                        const int* __restrict__ flag) {
  double ib;
  if (*flag) {  // uniform
    const double bn = sqrt(*nrm2);
    ib = bn > 1e-300 ? 1.0 / bn : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      beta[j] = bn;
      beta[kMaxBasis + j] = ib;
      if (j + 1 < m) {
        T[size_t(j) * m + j + 1] = bn;
        T[size_t(j + 1) * m + j] = bn;
      }
    }
  } else {
    ib = beta[kMaxBasis + j];
  }
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    v[p] = float(double(w[p]) * ib);
}

// out[i] = sum_j Y[j][i] V[j], i < nk, j < m <= kRitzMaxM (FP64
// accumulation), kRitzCols outputs per pass over V.  Y is read with
// wave-uniform indices, so its values come through the scalar cache and feed
// the FMAs as scalar operands; the basis is loaded 16 rows at a time, all 16
// loads issued before use (memory-level parallelism bounds this kernel).
constexpr int kRitzMaxM = 344;
constexpr int kRitzCols = 24;

__global__ __launch_bounds__(256) void ritz_kernel(const float* __restrict__ V, size_t ld, int m, int nk,
                                                   const double* __restrict__ Y /* m x nk */, float* __restrict__ out,
                                                   size_t Np) {
  for (int i0 = 0; i0 < nk; i0 += kRitzCols) {
    const int cnt = nk - i0 < kRitzCols ? nk - i0 : kRitzCols;
    for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
      double acc[kRitzCols];
#pragma unroll
      for (int i = 0; i < kRitzCols; i++) acc[i] = 0;
      for (int j0 = 0; j0 < m; j0 += 16) {
        float v[16];
#pragma unroll
        for (int t = 0; t < 16; t++) v[t] = V[size_t(j0 + t < m ? j0 + t : 0) * ld + p];
#pragma unroll
        for (int t = 0; t < 16; t++) {
          if (j0 + t >= m) break;  // uniform
          const double* y = Y + size_t(j0 + t) * nk + i0;
          const double vt = double(v[t]);
#pragma unroll
          for (int i = 0; i < kRitzCols; i++)
            if (i < cnt) acc[i] += y[i] * vt;
        }
      }
#pragma unroll
      for (int i = 0; i < kRitzCols; i++)
        if (i < cnt) out[size_t(i0 + i) * ld + p] = float(acc[i]);
    }
  }
}

// out_c = sum_j Y[j][c] V_j for c < nk with Y staged in LDS, zero-padded to
// ncp = nk rounded up to NC columns, so that each row's NC coefficients are
// read unconditionally as 16-byte pairs.  LDS bandwidth bounds a broadcast
// read per FMA, so each coefficient pair read serves PX pixels per thread
// (pixels p0 + k * stride, coalesced per k); the loads of the next 4 basis
// rows x PX pixels are issued before the current rows' FMAs where registers
// allow (NC * PX <= 64).  One pass over V per NC columns.
template <int NC, int PX>
__global__ __launch_bounds__(256) void ritz_lds_kernel(const float* __restrict__ V, size_t ld, int m, int nk,
                                                       const double* __restrict__ Y, float* __restrict__ out,
                                                       size_t ldo, size_t Np) {
  extern __shared__ double ys[];  // [m][ncp]
  const int ncp = (nk + NC - 1) / NC * NC;
  for (int t = threadIdx.x; t < m * ncp; t += blockDim.x) {
    const int j = t / ncp, c = t - j * ncp;
    ys[t] = c < nk ? Y[size_t(j) * nk + c] : 0.0;
  }
  __syncthreads();
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (int c0 = 0; c0 < nk; c0 += NC) {
    for (size_t p0 = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p0 < Np; p0 += stride * PX) {
      size_t pp[PX];
#pragma unroll
      for (int k = 0; k < PX; k++) {
        const size_t p = p0 + size_t(k) * stride;
        pp[k] = p < Np ? p : Np - 1;
      }
      double acc[PX][NC];
#pragma unroll
      for (int k = 0; k < PX; k++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[k][c] = 0.0;
      // with kPrefetch, rows j0 .. j0+3 are consumed while rows j0+4 .. j0+7
      // load (only where the extra registers keep the occupancy)
      constexpr bool kPrefetch = NC * PX <= 64;
      float v[4][PX], vn[4][PX];
      if constexpr (kPrefetch) {
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
          for (int k = 0; k < PX; k++) vn[t][k] = V[size_t(t < m ? t : m - 1) * ld + pp[k]];
      }
      for (int j0 = 0; j0 < m; j0 += 4) {
        if constexpr (kPrefetch) {
#pragma unroll
          for (int t = 0; t < 4; t++)
#pragma unroll
            for (int k = 0; k < PX; k++) v[t][k] = vn[t][k];
          if (j0 + 4 < m) {  // uniform
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
              for (int k = 0; k < PX; k++) vn[t][k] = V[size_t(j0 + 4 + t < m ? j0 + 4 + t : m - 1) * ld + pp[k]];
          }
        } else {
#pragma unroll
          for (int t = 0; t < 4; t++)
#pragma unroll
            for (int k = 0; k < PX; k++) v[t][k] = V[size_t(j0 + t < m ? j0 + t : m - 1) * ld + pp[k]];
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
          if (j0 + t >= m) break;  // uniform
          const double2* yr = reinterpret_cast<const double2*>(ys + size_t(j0 + t) * ncp + c0);
#pragma unroll
          for (int c2 = 0; c2 < NC / 2; c2++) {
            const double2 y2 = yr[c2];
#pragma unroll
            for (int k = 0; k < PX; k++) {
              acc[k][2 * c2] = fma(y2.x, double(v[t][k]), acc[k][2 * c2]);
              acc[k][2 * c2 + 1] = fma(y2.y, double(v[t][k]), acc[k][2 * c2 + 1]);
            }
          }
        }
      }
#pragma unroll
      for (int k = 0; k < PX; k++) {
        if (p0 + size_t(k) * stride >= Np) continue;
#pragma unroll
        for (int c = 0; c < NC; c++)
          if (c0 + c < nk) out[size_t(c0 + c) * ldo + pp[k]] = float(acc[k][c]);
      }
    }
  }
}

// Interior-only helpers
__global__ __launch_bounds__(256) void sqrt_degree_kernel(const float* __restrict__ D, float* __restrict__ v, size_t Np) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    v[p] = D[p] > 0.f ? sqrtf(D[p]) : 0.f;
}

__global__ __launch_bounds__(256) void random_start_kernel(float* __restrict__ v, int W, int H, int Wp, size_t Np,
                                                           uint64_t seed) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    const int y = int(p / Wp) - kICRadius, x = int(p % Wp) - kICRadius;
    float r = 0.f;
    if (x >= 0 && x < W && y >= 0 && y < H) {
      uint64_t z = seed + p * 0x9E3779B97F4A7C15ull;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      r = float(double(z >> 11) * (1.0 / 9007199254740992.0) - 0.5);
    }
    v[p] = r;
  }
}

// v = D^1/2 sum_{i=1}^{n-1} e_i: the u-space image of approximate generalised
// eigenvectors e_i (the trivial e_0 is left out: the solver locks D^1/2 1).
__global__ __launch_bounds__(256) void combine_start_kernel(const float* __restrict__ ev, int n,
                                                            const float* __restrict__ D, float* __restrict__ v,
                                                            size_t Np) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    float a = 0.f;
    for (int i = 1; i < n; i++) a += ev[size_t(i) * Np + p];
    v[p] = D[p] > 0.f ? sqrtf(D[p]) * a : 0.f;
  }
}

__global__ __launch_bounds__(256) void scale_by_kernel(float* __restrict__ v, const double* __restrict__ nrm2, size_t Np) {
  const double ib = 1.0 / sqrt(nrm2[0]);
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    v[p] = float(double(v[p]) * ib);
}

// generalised eigenvector v = s .* u
__global__ __launch_bounds__(256) void to_general_kernel(const float* __restrict__ u, const float* __restrict__ s,
                                                         float* __restrict__ v, size_t Np) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    v[p] = u[p] * s[p];
}

// ---- launch helpers shared with the block solver -----------------------------
void launch_reduce(Context& c, const double* partial, int G, int n, double* dst) {
  hipLaunchKernelGGL(reduce_kernel, dim3(n), dim3(kRedThreads), 0, c.stream, partial, G, dst, nullptr);
  GPB_LAUNCH_CHECK();
}

void launch_reduce_gated(Context& c, const double* partial, int G, int n, double* dst, const int* gate) {
  hipLaunchKernelGGL(reduce_kernel, dim3(n), dim3(kRedThreads), 0, c.stream, partial, G, dst, gate);
  GPB_LAUNCH_CHECK();
}

void launch_ritz(Context& c, const float* V, size_t ld, int m, int nk, const double* Ydev, float* out, size_t Np) {
  if (m > kRitzMaxM) throw ArgError("basis too large for the Ritz recombination");
  const dim3 grid(std::min<unsigned>(2048, div_up(Np, 256)));
  const int nc = nk <= 16 ? 16 : (nk <= 24 ? 24 : 32);
  const size_t lds = sizeof(double) * size_t(m) * size_t((nk + nc - 1) / nc * nc);
  if (lds <= 64 * 1024) {
    KScope ks(c, kKRitz, double(Np) * 4.0 * (double(m) * div_up(nk, nc) + nk));
    if (nc == 16)
      hipLaunchKernelGGL((ritz_lds_kernel<16, 4>), grid, dim3(256), lds, c.stream, V, ld, m, nk, Ydev, out, ld, Np);
    else if (nc == 24)
      hipLaunchKernelGGL((ritz_lds_kernel<24, 2>), grid, dim3(256), lds, c.stream, V, ld, m, nk, Ydev, out, ld, Np);
    else
      hipLaunchKernelGGL((ritz_lds_kernel<32, 2>), grid, dim3(256), lds, c.stream, V, ld, m, nk, Ydev, out, ld, Np);
    GPB_LAUNCH_CHECK();
    return;
  }
  KScope ks(c, kKRitz, double(Np) * 4.0 * (m + nk));
  hipLaunchKernelGGL(ritz_kernel, grid, dim3(256), 0, c.stream, V, ld, m, nk, Ydev, out, Np);
  GPB_LAUNCH_CHECK();
}

void launch_locked_start(Context& c, const Stencil& st, float* v0, float* v1_block, int B, size_t ld, uint64_t seed) {
  const unsigned vg = std::min<unsigned>(2048, div_up(st.Np, 256));
  hipLaunchKernelGGL(sqrt_degree_kernel, dim3(vg), dim3(256), 0, c.stream, st.D, v0, st.Np);
  GPB_LAUNCH_CHECK();
  for (int r = 0; r < B; r++) {
    hipLaunchKernelGGL(random_start_kernel, dim3(vg), dim3(256), 0, c.stream, v1_block + size_t(r) * ld, st.W, st.H,
                       st.Wp, st.Np, seed + 0x9E3779B97F4A7C15ull * uint64_t(r + 1));
    GPB_LAUNCH_CHECK();
  }
}

void launch_to_general(Context& c, const float* u, const float* s, float* v, size_t Np) {
  hipLaunchKernelGGL(to_general_kernel, dim3(std::min<unsigned>(2048, div_up(Np, 256))), dim3(256), 0, c.stream, u, s, v,
                     Np);
  GPB_LAUNCH_CHECK();
}

void launch_scale_by(Context& c, float* v, const double* nrm2, size_t Np) {
  hipLaunchKernelGGL(scale_by_kernel, dim3(std::min<unsigned>(2048, div_up(Np, 256))), dim3(256), 0, c.stream, v, nrm2,
                     Np);
  GPB_LAUNCH_CHECK();
}

namespace {

struct Reducer {
  Context& c;
  size_t Np;
  int NW;   // element chunks of the dot kernel
  int GU;   // blocks of the update kernel
  int GP;   // blocks of the projection kernel
  double* partial;
  double* npart;
  Reducer(Context& cc, size_t np) : c(cc), Np(np) {
    NW = int((np + kDotChunk - 1) / kDotChunk);
    GU = int(std::min<size_t>(2048, std::max<size_t>(1, (np + 255) / 256)));
    GP = int(std::min<size_t>(512, std::max<size_t>(1, (np + 255) / 256)));
    partial = c.ws<double>("eig.partial", std::max(size_t(kMaxBasis + 2) * NW, size_t(kProjectMaxN + 2) * GP));
    npart = c.ws<double>("eig.npart", size_t(GU));
  }
  // dst[0..n) = V_i . w, dst[n] = w . (D) w if wsq
  void dots(const float* V, size_t ld, int n, const float* w, const float* D, bool wsq, double* dst) {
    {
      KScope ks(c, kKDots, double(Np) * (4.0 * n + 4 + (wsq && D ? 4 : 0)));
      hipLaunchKernelGGL(multidot2_kernel, dim3(NW, n + (wsq ? 1 : 0)), dim3(256), 0, c.stream, V, ld, n, w, D, Np,
                         partial);
      GPB_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(reduce_kernel, dim3(n + (wsq ? 1 : 0)), dim3(kRedThreads), 0, c.stream, partial, NW, dst, nullptr);
    GPB_LAUNCH_CHECK();
  }
  // [alpha: w <- w - alpha v_j - beta_{j-1} v_{j-1}]; dst[0..n) = V_i . w,
  // dst[n] = ||w||^2 (n <= kProjectMaxN); gated when gate is non-null.
  void project(const float* V, size_t ld, int n, float* w, const double* alpha, const double* beta, int j,
               const int* gate, double* dst) {
    {
      KScope ks(c, kKDots, gate ? 0.0 : double(Np) * (4.0 * n + 4 + (alpha ? 12 : 0)));
      launch_project(c, V, ld, n, w, alpha, beta, j, gate, Np, partial, GP);
    }
    hipLaunchKernelGGL(reduce_kernel, dim3(n + 1), dim3(kRedThreads), 0, c.stream, partial, GP, dst, gate);
    GPB_LAUNCH_CHECK();
  }
  // dst[0..n) = V_i . w, dst[n] = ||w||^2 for any n < kMaxBasis, in chunks of
  // kProjectMaxN vectors (each chunk's norm slot is overwritten by the next
  // chunk's first dot; the last chunk's stays); gated when gate is non-null.
  void project_any(const float* V, size_t ld, int n, float* w, const int* gate, double* dst) {
    for (int i0 = 0; i0 < n; i0 += kProjectMaxN)
      project(V + size_t(i0) * ld, ld, std::min(kProjectMaxN, n - i0), w, nullptr, nullptr, 0, gate, dst + i0);
  }
  // w -= V h, then dst[0..n) = V_i . w, dst[n] = ||w||^2 of the new w (n <=
  // kProjectMaxN): the update and the next projection of classical
  // Gram-Schmidt in one pass, with the separate kernels' results
  void update_project(const float* V, size_t ld, int n, const double* h, float* w, double* dst) {
    {
      KScope ks(c, kKUpdate, double(Np) * (4.0 * n + 8));
      launch_update_project(c, V, ld, n, h, w, Np, partial, GP);
    }
    hipLaunchKernelGGL(reduce_kernel, dim3(n + 1), dim3(kRedThreads), 0, c.stream, partial, GP, dst, nullptr);
    GPB_LAUNCH_CHECK();
  }
  // dst[k] = v_k . D v_k for the vectors V + sel[k] ld, k < m, in one launch
  void wnorms(const float* V, size_t ld, const int* sel, int m, const float* D, double* dst) {
    if (m < 1 || m > kMaxNvec + 2) throw ArgError("too many norms");
    NormSel ns{};
    for (int k = 0; k < m; k++) ns.idx[k] = sel[k];
    {
      KScope ks(c, kKDots, double(Np) * (4.0 * m + 4));
      hipLaunchKernelGGL(wnorms_kernel, dim3(NW, m), dim3(256), 0, c.stream, V, ld, ns, D, Np, partial);
      GPB_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(reduce_kernel, dim3(m), dim3(kRedThreads), 0, c.stream, partial, NW, dst, nullptr);
    GPB_LAUNCH_CHECK();
  }
  // w -= V h; nrm_dst = ||w||^2 afterwards (if non-null); gated when gate is non-null
  void update(const float* V, size_t ld, int n, const double* h, float* w, double* nrm_dst,
              const int* gate = nullptr) {
    {
      KScope ks(c, kKUpdate, gate ? 0.0 : double(Np) * (4.0 * n + 8));
      if (n <= kProjectMaxN) {
        launch_update(c, V, ld, n, h, w, Np, nrm_dst ? npart : nullptr, gate, GU);
      } else {
        hipLaunchKernelGGL(update_wide_kernel, dim3(GU), dim3(256), 0, c.stream, V, ld, n, h, w, Np,
                           nrm_dst ? npart : nullptr, gate);
        GPB_LAUNCH_CHECK();
      }
    }
    if (nrm_dst) {
      hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kRedThreads), 0, c.stream, npart, GU, nrm_dst, gate);
      GPB_LAUNCH_CHECK();
    }
  }
};

unsigned vec_grid(size_t Np) { return std::min<unsigned>(2048, div_up(Np, 256)); }

}  // namespace

void lanczos_solve(Context& c, const Stencil& st, const EigOptions& opt_in, double* evals, float* d_evecs,
                   LanczosStats& stats, const float* start_evecs, int start_count) {
  EigOptions opt = opt_in;
  const size_t Np = st.Np;
  const size_t N = size_t(st.W) * st.H;
  if (opt.K < 2) throw ArgError("need at least 2 eigenpairs");
  if (N < size_t(opt.K) + 4) throw ArgError("image too small for the requested eigenpairs");
  opt.m = std::min<int>(std::min<int>(opt.m, kMaxBasis - 1), int(N) - 1);
  if (opt.m < opt.K + 2) throw ArgError("basis too small");
  opt.keep = std::max(opt.K + 1, std::min(opt.keep, opt.m - 2));
  const int m = opt.m, K = opt.K, keep = opt.keep;
  float* V = c.ws<float>("eig.V", size_t(m + 1) * Np);
  float* V2 = c.ws<float>("eig.V2", size_t(m + 1) * Np);
  float* w = c.ws<float>("eig.w", Np);
  double* T = c.ws<double>("eig.T", size_t(m) * m);
  double* h = c.ws<double>("eig.h", kMaxBasis + 2);
  double* h1 = c.ws<double>("eig.h1", kMaxBasis + 2);
  double* h2 = c.ws<double>("eig.h2", kMaxBasis + 2);
  double* nrm = c.ws<double>("eig.nrm", 2);
  int* flag = c.ws<int>("eig.flag", 1);
  double* gacc = c.ws<double>("eig.gacc", 3);
  GPB_HIP(hipMemsetAsync(gacc, 0, 3 * sizeof(double), c.stream));
  double gacc_h[3] = {0, 0, 0};
  int* one_flag = c.ws<int>("eig.one", 1);
  hipLaunchKernelGGL(fill_int1_kernel, dim3(1), dim3(1), 0, c.stream, one_flag, 1);
  GPB_LAUNCH_CHECK();
  double* beta = c.ws<double>("eig.beta", 2 * kMaxBasis);
  double* alpha = c.ws<double>("eig.alpha", 1);
  const int spmv_g = spmv_blocks(st, tune_spmv(c, st));
  const int step_g = (c.reorth_mode >= 1 && eig_partial_reorth(opt) && c.lanczos_fused)
                         ? spmv_blocks(st, tune_lanczos_step(c, st))
                         : 1;
  double* xy_part = c.ws<double>("eig.xy", size_t(spmv_g));
  double* Ydev = c.ws<double>("eig.Y", size_t(m) * m);
  // partial reorthogonalisation for targets the FP64 refinement does not
  // follow (see kRefineBelowTol); whole-basis Gram-Schmidt every step otherwise
  const bool pro = c.reorth_mode >= 1 && eig_partial_reorth(opt);
  // optimistic cycles (mode 2): no gated launches; a cycle whose omega
  // recurrence triggered is redone from that step with them (below)
  const bool optimistic = pro && c.reorth_mode == 2;
  const double pro_thr = c.pro_threshold > 0 ? c.pro_threshold : kProSqrtEps;
  // fused steps (gpb_set_lanczos_fusion): one pass per three-term step (the
  // SpMV with the previous step's scaling and this step's update around it,
  // lanczos_step_tile) and one small kernel for alpha, beta and omega.  Their
  // w alternates between two buffers (neighbouring tiles read the previous w
  // in their halos while a tile writes the new one), and they write only
  // interior pixels, so every basis halo starts at zero.
  const bool fused = pro && c.lanczos_fused;
  float* wbuf[2] = {w, fused ? c.ws<float>("eig.w2", Np) : w};
  auto wj = [&](int j) { return wbuf[j & 1]; };
  double* aeff = c.ws<double>("eig.aeff", kMaxBasis);
  double* alpha_f = c.ws<double>("eig.alpha_f", 1);
  double* part3 = c.ws<double>("eig.part3", size_t(3) * step_g);
  if (fused) {
    launch_fill(c, wbuf[1], Np, 0.f);
    launch_fill(c, V, size_t(m + 1) * Np, 0.f);
    launch_fill(c, V2, size_t(m + 1) * Np, 0.f);
  }
  // one omega vector per step, so that a cycle can be resumed at any step
  double* om = c.ws<double>("eig.omega", size_t(kMaxBasis + 2) * (kMaxBasis + 1));
  int* first_viol = c.ws<int>("eig.first_viol", 1);
  int* zero_flag = c.ws<int>("eig.zero", 1);
  hipLaunchKernelGGL(fill_int1_kernel, dim3(1), dim3(1), 0, c.stream, zero_flag, 0);
  GPB_LAUNCH_CHECK();
  int* pend = c.ws<int>("eig.pending", 1);
  double* rnrm = c.ws<double>("eig.rnrm", 1);
  auto om_slot = [&](int i) { return om + size_t(i) * (kMaxBasis + 1); };
  Reducer red(c, Np);
  const unsigned vg = vec_grid(Np);
  launch_fill(c, w, Np, 0.f);
  GPB_HIP(hipMemsetAsync(T, 0, sizeof(double) * m * m, c.stream));
  // V0 = D^1/2 1 / norm
  hipLaunchKernelGGL(sqrt_degree_kernel, dim3(vg), dim3(256), 0, c.stream, st.D, V, Np);
  GPB_LAUNCH_CHECK();
  red.dots(V, Np, 0, V, nullptr, true, h);
  hipLaunchKernelGGL(scale_by_kernel, dim3(vg), dim3(256), 0, c.stream, V, h, Np);
  GPB_LAUNCH_CHECK();
  // T00 = V0' L V0
  launch_spmv(c, st, V, w);
  red.dots(V, Np, 1, w, nullptr, false, h);
  hipLaunchKernelGGL(tcol_kernel, dim3(1), dim3(64), 0, c.stream, T, m, 0, 0, nullptr, h, h, nullptr, h, h, nullptr);
  GPB_LAUNCH_CHECK();
  // V1: random (or the given vectors' combination), orthogonal to V0, normalised
  float* V1 = V + Np;
  if (start_evecs && start_count > 1)
    hipLaunchKernelGGL(combine_start_kernel, dim3(vg), dim3(256), 0, c.stream, start_evecs, start_count, st.D, V1, Np);
  else
    hipLaunchKernelGGL(random_start_kernel, dim3(vg), dim3(256), 0, c.stream, V1, st.W, st.H, st.Wp, Np,
                       uint64_t(0x5eed1234abcdull));
  GPB_LAUNCH_CHECK();
  for (int pass = 0; pass < 2; pass++) {
    red.dots(V, Np, 1, V1, nullptr, false, h);
    red.update(V, Np, 1, h, V1, nullptr);
  }
  red.dots(V, Np, 0, V1, nullptr, true, h);
  hipLaunchKernelGGL(scale_by_kernel, dim3(vg), dim3(256), 0, c.stream, V1, h, Np);
  GPB_LAUNCH_CHECK();

  int k = 1;
  stats = LanczosStats();
  stats.basis = m;
  stats.block = 1;
  int matvecs = 1;
  std::vector<double> th, Y;
  // restart exchanges go through pinned host buffers (asynchronous DMA; the
  // host reads them after a stream synchronisation and rewrites the upload
  // buffers only after the next one, by which time their copies are done)
  double* Th = c.host_ws<double>("eig.h.T", size_t(m) * m);
  double* betah = c.host_ws<double>("eig.h.beta", kMaxBasis);
  int* viol_h = c.host_ws<int>("eig.h.viol", 1);
  double* Yk = c.host_ws<double>("eig.h.Y", size_t(m) * keep);
  double* wnrm_h = c.host_ws<double>("eig.h.nrm", K + 1);
  double* gacc_p = c.host_ws<double>("eig.h.gacc", 3);
  double* T0 = c.host_ws<double>("eig.h.T0", size_t(m) * m);
  bool converged = false;
  std::vector<double> resid(K, 0.0);
  // One restart cycle's device work, steps k0 .. m-1 on the basis Vc.  From
  // the first restart on (k0 = keep) the launch sequence depends only on
  // which of V / V2 is current, so it is captured once per buffer into a
  // hipGraph and replayed; the first cycle, and every cycle while per-kernel
  // timing is on, launches directly.
  // j0 > k0 resumes a cycle at step j0 (its steps before j0 already done),
  // launching the gated whole-basis pass on every step.
  auto cycle = [&](int k0, float* Vc, int j0, bool opt_cycle) {
    float* const V = Vc;
    const int k = k0;
    if (opt_cycle) {
      hipLaunchKernelGGL(fill_int1_kernel, dim3(1), dim3(1), 0, c.stream, first_viol, kMaxBasis + 1);
      GPB_LAUNCH_CHECK();
    }
    for (int j = j0; j < m; j++) {
      float* vj = V + size_t(j) * Np;
      float* const w = wj(j);
      const int n = j + 1;
      if (j == k || (!pro && m > kProjectMaxN)) {
        // first step of a cycle (its column couples to every kept Ritz
        // vector): classical Gram-Schmidt twice, FP64 accumulation
        launch_spmv(c, st, vj, w);
        if (n <= kProjectMaxN) {
          red.project(V, Np, n, w, nullptr, nullptr, j, nullptr, h1);
          red.update_project(V, Np, n, h1, w, h2);
        } else {
          red.dots(V, Np, n, w, nullptr, false, h1);
          red.update(V, Np, n, h1, w, nullptr);
          red.dots(V, Np, n, w, nullptr, false, h2);
        }
        red.update(V, Np, n, h2, w, nrm + 1);
        hipLaunchKernelGGL(tcol_kernel, dim3(1), dim3(kMaxBasis), 0, c.stream, T, m, j, int(pro || m <= kProjectMaxN),
                           nullptr,
                           h1, h2, one_flag,
                           nrm + 1, nrm + 1, beta);
        GPB_LAUNCH_CHECK();
        if (pro) {
          hipLaunchKernelGGL(omega_reset_kernel, dim3(1), dim3(256), 0, c.stream, om_slot(j), om_slot(j + 1), j, pend,
                             aeff);
          GPB_LAUNCH_CHECK();
        }
      } else if (fused) {
        // v_j formed from the previous step's w while the tile is staged --
        // or, for the first step of a resumed cycle, read from the basis,
        // which holds it -- then w = L v_j - beta_{j-1} v_{j-1} and alpha_j,
        // beta_j and omega from its partial sums
        const bool resumed = j == j0 && j0 > k;
        launch_lanczos_step(c, st, resumed ? 0 : 1, wj(j - 1), V + size_t(j - 1) * Np, vj, w, aeff, beta, j, part3);
        hipLaunchKernelGGL(omega_fused_kernel, dim3(1), dim3(256), 0, c.stream, T, m, j, part3, step_g, beta, aeff,
                           alpha_f, om_slot(j), om_slot(j - 1), om_slot(j + 1), flag, pend, double(Np), gacc,
                           pro_thr, opt_cycle ? first_viol : nullptr);
        GPB_LAUNCH_CHECK();
        if (!opt_cycle) {
          hipLaunchKernelGGL(materialise_kernel, dim3(vg), dim3(256), 0, c.stream, w, vj, alpha_f, Np, flag);
          GPB_LAUNCH_CHECK();
          red.project_any(V, Np, n, w, flag, h2);
          red.update(V, Np, n, h2, w, nrm + 1, flag);
          hipLaunchKernelGGL(gated_beta_kernel, dim3(1), dim3(1), 0, c.stream, T, m, j, beta, aeff, nrm + 1, flag);
          GPB_LAUNCH_CHECK();
        }
        continue;
      } else if (pro) {
        // three-term step; the whole-basis pass only when the omega
        // recurrence says semi-orthogonality is about to be lost
        launch_spmv(c, st, vj, w, xy_part);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kRedThreads), 0, c.stream, xy_part, spmv_g, alpha, nullptr);
        GPB_LAUNCH_CHECK();
        {
          KScope ks(c, kKUpdate, double(Np) * 16.0);
          hipLaunchKernelGGL(three_term_kernel, dim3(red.GU), dim3(256), 0, c.stream, V, Np, j, alpha, beta, w, Np,
                             red.npart);
          GPB_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(omega_kernel, dim3(1), dim3(256), 0, c.stream, T, m, j, alpha, beta, red.npart, red.GU,
                           om_slot(j), om_slot(j - 1), om_slot(j + 1), flag, pend, double(Np), gacc, pro_thr,
                           opt_cycle ? first_viol : nullptr);
        GPB_LAUNCH_CHECK();
        if (!opt_cycle) {
          red.project_any(V, Np, n, w, flag, h2);
          red.update(V, Np, n, h2, w, nrm + 1, flag);
        }
        hipLaunchKernelGGL(pro_scale_kernel, dim3(vg), dim3(256), 0, c.stream, w, beta, j, V + size_t(j + 1) * Np, Np, T,
                           m, nrm + 1, opt_cycle ? zero_flag : flag);
        GPB_LAUNCH_CHECK();
        continue;
      } else {
        // three-term step, one full projection pass, and a second pass only
        // when the DGKS test asks for it
        launch_spmv(c, st, vj, w, xy_part);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kRedThreads), 0, c.stream, xy_part, spmv_g, alpha, nullptr);
        GPB_LAUNCH_CHECK();
        red.project(V, Np, n, w, alpha, beta, j, nullptr, h1);  // h1[n] = ||w'||^2
        red.update(V, Np, n, h1, w, nrm);
        hipLaunchKernelGGL(reorth_decide_kernel, dim3(1), dim3(1), 0, c.stream, h1 + n, nrm, flag, n, double(Np), gacc);
        GPB_LAUNCH_CHECK();
        red.project(V, Np, n, w, nullptr, nullptr, j, flag, h2);
        red.update(V, Np, n, h2, w, nrm + 1, flag);
        hipLaunchKernelGGL(tcol_kernel, dim3(1), dim3(kMaxBasis), 0, c.stream, T, m, j, 1, alpha, h1, h2, flag, nrm,
                           nrm + 1, beta);
        GPB_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(scale_kernel, dim3(vg), dim3(256), 0, c.stream, w, beta, j, V + size_t(j + 1) * Np, Np);
      GPB_LAUNCH_CHECK();
    }
    if (fused) {
      // the residual direction v_m: the scaling a next step would have done
      hipLaunchKernelGGL(fused_last_kernel, dim3(vg), dim3(256), 0, c.stream, wj(m - 1), V + size_t(m - 1) * Np, aeff,
                         beta, m - 1, V + size_t(m) * Np, Np);
      GPB_LAUNCH_CHECK();
    }
  };
  const bool use_graph = c.lanczos_graph && !c.kt.on;
  hipGraphExec_t graphs[2] = {nullptr, nullptr};
  int parity = 0;  // which of the two basis buffers is current
  struct GraphGuard {
    hipGraphExec_t* g;
    ~GraphGuard() {
      for (int i = 0; i < 2; i++)
        if (g[i]) (void)hipGraphExecDestroy(g[i]);
    }
  } graph_guard{graphs};
  // The kept Ritz vectors inherit the basis' loss of orthogonality, and the
  // residual vector's overlap with them would feed every three-term step of
  // the next cycle (the omega recurrence assumes it is at rounding level):
  // Gram-Schmidt it against them twice, then normalise.
  auto clean_residual = [&](float* Vn) {
    float* vk = Vn + size_t(keep) * Np;
    if (keep <= kProjectMaxN) {
      red.project(Vn, Np, keep, vk, nullptr, nullptr, 0, nullptr, h1);
      red.update_project(Vn, Np, keep, h1, vk, h2);
      red.update(Vn, Np, keep, h2, vk, rnrm);
    } else {
      for (int pass = 0; pass < 2; pass++) {
        red.project_any(Vn, Np, keep, vk, nullptr, h1);
        red.update(Vn, Np, keep, h1, vk, rnrm);
      }
    }
    launch_scale_by(c, vk, rnrm, Np);
  };
  // With partial reorthogonalisation the clean-up depends only on the new
  // basis, so it is queued before the host reads the residual norms and runs
  // while the host decides (after the last restart it is work on a vector no
  // output uses).  Not under per-kernel timing, whose events are read at that
  // point.
  const bool early_clean = pro && !c.kt.on;
  hipEvent_t norms_ready = nullptr;
  if (early_clean) GPB_HIP(hipEventCreateWithFlags(&norms_ready, hipEventDisableTiming));
  struct EventGuard {
    hipEvent_t e;
    ~EventGuard() {
      if (e) (void)hipEventDestroy(e);
    }
  } norms_guard{norms_ready};
  while (true) {
    if (use_graph && k == keep) {
      hipGraphExec_t& ge = graphs[parity];
      if (!ge) {
        hipGraph_t g;
        GPB_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        c.capturing = true;
        try {
          cycle(k, V, k, optimistic);
        } catch (...) {
          c.capturing = false;
          hipGraph_t dummy;
          (void)hipStreamEndCapture(c.stream, &dummy);
          throw;
        }
        c.capturing = false;
        GPB_HIP(hipStreamEndCapture(c.stream, &g));
        GPB_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        GPB_HIP(hipGraphDestroy(g));
      }
      GPB_HIP(hipGraphLaunch(ge, c.stream));
    } else {
      cycle(k, V, k, optimistic);
    }
    matvecs += m - k;
    // Rayleigh-Ritz on the host
    *viol_h = kMaxBasis + 1;
    GPB_HIP(hipMemcpyAsync(Th, T, sizeof(double) * m * m, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipMemcpyAsync(betah, beta, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
    if (optimistic) GPB_HIP(hipMemcpyAsync(viol_h, first_viol, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    if (optimistic && *viol_h < m) {
      // the omega recurrence triggered at step viol_h: every step before it
      // matches the gated form, so redo the cycle from there with the gated
      // pass.  The steps after it are discarded; of what they leave behind,
      // the twice-rule flag is cleared (the gated form has it clear at that
      // step) and so is w: the SpMV writes interior pixels only, so w's zero
      // halo must not carry their values (unnormalisable vectors there scale
      // to inf and NaN).  Basis vectors up to that step predate it, later ones
      // are rewritten whole.
      hipLaunchKernelGGL(fill_int1_kernel, dim3(1), dim3(1), 0, c.stream, pend, 0);
      GPB_LAUNCH_CHECK();
      launch_fill(c, w, Np, 0.f);
      const int redo = *viol_h;
      cycle(k, V, redo, false);
      matvecs += m - redo;
      stats.redone_steps += m - redo;
      GPB_HIP(hipMemcpyAsync(Th, T, sizeof(double) * m * m, hipMemcpyDeviceToHost, c.stream));
      GPB_HIP(hipMemcpyAsync(betah, beta, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
      GPB_HIP(hipStreamSynchronize(c.stream));
    }
    std::vector<double> A(Th, Th + size_t(m) * m);
    {
      // smallest Ritz values first
      std::vector<double> thd, Yd;
      symeig_desc(m, A, thd, Yd);
      th.assign(size_t(m), 0.0);
      Y.assign(size_t(m) * m, 0.0);
      for (int i = 0; i < m; i++) {
        th[size_t(i)] = thd[size_t(m - 1 - i)];
        for (int j = 0; j < m; j++) Y[size_t(j) * m + i] = Yd[size_t(j) * m + (m - 1 - i)];
      }
    }
    const int nk = keep;
    for (int j = 0; j < m; j++)
      for (int i = 0; i < nk; i++) Yk[size_t(j) * nk + i] = Y[size_t(j) * m + i];
    GPB_HIP(hipMemcpyAsync(Ydev, Yk, sizeof(double) * m * nk, hipMemcpyHostToDevice, c.stream));
    launch_ritz(c, V, Np, m, nk, Ydev, V2, Np);
    GPB_HIP(hipMemcpyAsync(V2 + size_t(nk) * Np, V + size_t(m) * Np, Np * sizeof(float), hipMemcpyDeviceToDevice,
                           c.stream));
    // weighted norms ||D^1/2 u_i||, ||D^1/2 v_m|| in one launch
    {
      int sel[kMaxNvec + 2];
      for (int i = 0; i < K; i++) sel[i] = i;
      sel[K] = nk;
      red.wnorms(V2, Np, sel, K + 1, st.D, h);
    }
    GPB_HIP(hipMemcpyAsync(wnrm_h, h, sizeof(double) * (K + 1), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipMemcpyAsync(gacc_p, gacc, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (early_clean) {
      GPB_HIP(hipEventRecord(norms_ready, c.stream));
      clean_residual(V2);
      GPB_HIP(hipEventSynchronize(norms_ready));
    } else {
      GPB_HIP(hipStreamSynchronize(c.stream));
    }
    std::copy(gacc_p, gacc_p + 3, gacc_h);
    c.kt.fold();
    if (c.kt.timing(kKDots)) c.kt.bytes[kKDots] += gacc_h[0] - stats.pass2_bytes[0];
    if (c.kt.timing(kKUpdate)) c.kt.bytes[kKUpdate] += gacc_h[1] - stats.pass2_bytes[1];
    stats.pass2_bytes[0] = gacc_h[0];
    stats.pass2_bytes[1] = gacc_h[1];
    stats.second_passes = int(gacc_h[2]);
    std::swap(V, V2);
    parity ^= 1;
    const double bm = betah[m - 1];
    double worst = 0;
    for (int i = 0; i < K; i++) {
      resid[i] = std::fabs(bm * Y[size_t(m - 1) * m + i]) * std::sqrt(wnrm_h[K]) / std::sqrt(std::max(wnrm_h[i], 1e-300));
      worst = std::max(worst, resid[i]);
    }
    stats.max_residual = worst;
    if (c.verbose)
      std::fprintf(stderr, "gpb: Lanczos %dx%d: %d operator applications, residual %.3e\n", st.W, st.H, matvecs, worst);
    converged = conv_count(opt, th.data(), resid.data()) == K;
    if (converged || matvecs >= opt.max_matvecs) break;
    // thick restart: basis = Ritz vectors [0, keep) and v_m; T = diag(theta)
    std::fill(T0, T0 + size_t(m) * m, 0.0);
    for (int i = 0; i < keep; i++) T0[size_t(i) * m + i] = th[size_t(i)];
    GPB_HIP(hipMemcpyAsync(T, T0, sizeof(double) * m * m, hipMemcpyHostToDevice, c.stream));
    k = keep;
    stats.restarts++;
    if (pro && !early_clean) clean_residual(V);
  }
  stats.iterations = matvecs;
  stats.converged = 0;
  stats.converged = th.empty() ? 0 : conv_count(opt, th.data(), resid.data());
  for (int i = 0; i < K; i++) {
    evals[i] = th[size_t(i)];
    hipLaunchKernelGGL(to_general_kernel, dim3(vg), dim3(256), 0, c.stream, V + size_t(i) * Np, st.s,
                       d_evecs + size_t(i) * Np, Np);
    GPB_LAUNCH_CHECK();
  }
  if (!converged)
    throw NoConvergence("Lanczos stopped after " + std::to_string(matvecs) + " operator applications with residual " +
                        std::to_string(stats.max_residual));
}

}  // namespace gpb

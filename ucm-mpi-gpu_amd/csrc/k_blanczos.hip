// Stage (d), block form: block thick-restart Lanczos on the normalised
// Laplacian L = D^-1/2 (D - W) D^-1/2 with B columns per step.
//
// Reference: the single-vector solvers of spectralPb.cpp:221-244 (Spectra)
// and damascene/src/noReorthog/lanczos.cu:393-752.  Each block step applies
// the stencil once to B vectors (W is read once for all of them), runs two
// classical Gram-Schmidt passes of the new panel against the basis as FP64
// panel products (the basis is read once per pass for all B columns), and
// orthonormalises the panel by Cholesky QR done twice (CholQR2), with the
// B x B Cholesky factorisation in a one-thread kernel so that a restart
// cycle needs no host synchronisation.  The vector D^1/2 1 (lambda = 0) is
// locked as basis vector 0.  The starting block is random or supplied by the
// caller (warm start).
#include <cmath>
#include <cstring>
#include <vector>

#include "gpb_pipeline.hpp"

namespace gpb {

void symeig_desc(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& z);

constexpr int kBMax = 16;
constexpr int kBTX = 32, kBTY = 8;
constexpr int kBTileW = kBTX + 2 * kICRadius, kBTileH = kBTY + 2 * kICRadius;
constexpr int kBDotChunk = 4096;
constexpr int kBMaxBasis = 336;

struct BSpmvArgs {
  int o[kStencilHalf];
  int lo[kStencilHalf];
};

// Y_r = L X_r for r < B; X, Y columns are ldx / ldy apart (padded layout).
template <int B>
__global__ __launch_bounds__(kBTX * kBTY) void bspmv_kernel(const float* __restrict__ wup, const float* __restrict__ s,
                                                            const float* __restrict__ X, size_t ldx,
                                                            float* __restrict__ Y, size_t ldy, int W, int H, int Wp,
                                                            int Hp, size_t Np, BSpmvArgs A) {
  __shared__ float u[kBTileH * kBTileW * B];
  const int tid = threadIdx.y * kBTX + threadIdx.x;
  const int x0 = blockIdx.x * kBTX, y0 = blockIdx.y * kBTY;
  for (int i = tid; i < kBTileH * kBTileW; i += kBTX * kBTY) {
    const int ty = i / kBTileW, tx = i - ty * kBTileW;
    const int py = y0 + ty, px = x0 + tx;
    if (py < Hp && px < Wp) {
      const size_t q = size_t(py) * Wp + px;
      const float sq = s[q];
#pragma unroll
      for (int r = 0; r < B; r++) u[i * B + r] = sq * X[size_t(r) * ldx + q];
    } else {
#pragma unroll
      for (int r = 0; r < B; r++) u[i * B + r] = 0.f;
    }
  }
  __syncthreads();
  const int xi = x0 + threadIdx.x, yi = y0 + threadIdx.y;
  if (xi >= W || yi >= H) return;
  const size_t p = size_t(yi + kICRadius) * Wp + xi + kICRadius;
  const int c = (threadIdx.y + kICRadius) * kBTileW + threadIdx.x + kICRadius;
  float up[B], acc[B];
#pragma unroll
  for (int r = 0; r < B; r++) {
    up[r] = u[c * B + r];
    acc[r] = 0.f;
  }
#pragma unroll 4
  for (int j = 0; j < kStencilHalf; j++) {
    const float w1 = wup[size_t(j) * Np + p], w2 = wup[size_t(j) * Np + p - A.o[j]];
    const float* u1 = u + (c + A.lo[j]) * B;
    const float* u2 = u + (c - A.lo[j]) * B;
#pragma unroll
    for (int r = 0; r < B; r++) {
      acc[r] = __builtin_fmaf(w1, up[r] - u1[r], acc[r]);
      acc[r] = __builtin_fmaf(w2, up[r] - u2[r], acc[r]);
    }
  }
  const float sp = s[p];
#pragma unroll
  for (int r = 0; r < B; r++) Y[size_t(r) * ldy + p] = sp * acc[r];
}

__device__ __forceinline__ double bwave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// partial[(i * B + r)][chunk] = sum over the chunk of V_i . (D) W_r, i < n.
template <int B>
__global__ __launch_bounds__(256) void bdots_kernel(const float* __restrict__ V, size_t ld, const float* __restrict__ Wm,
                                                    size_t ldw, const float* __restrict__ D, size_t Np,
                                                    double* __restrict__ partial) {
  __shared__ double sh[4][B];
  const int i = blockIdx.y;
  const size_t lo = size_t(blockIdx.x) * kBDotChunk;
  const float* v = V + size_t(i) * ld;
  double acc[B];
#pragma unroll
  for (int r = 0; r < B; r++) acc[r] = 0;
  for (int e = threadIdx.x; e < kBDotChunk; e += 256) {
    const size_t p = lo + e;
    if (p < Np) {
      double x = v[p];
      if (D) x *= double(D[p]);
#pragma unroll
      for (int r = 0; r < B; r++) acc[r] += x * double(Wm[size_t(r) * ldw + p]);
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < B; r++) {
    const double t = bwave_sum(acc[r]);
    if (lane == 0) sh[wv][r] = t;
  }
  __syncthreads();
  if (threadIdx.x < B) {
    const int r = threadIdx.x;
    partial[(size_t(i) * B + r) * gridDim.x + blockIdx.x] = sh[0][r] + sh[1][r] + sh[2][r] + sh[3][r];
  }
}

// W_r <- W_r - sum_i H[i][r] V_i (FP64 accumulation), grid-stride.
template <int B>
__global__ __launch_bounds__(256) void bupdate_kernel(const float* __restrict__ V, size_t ld, int n,
                                                      const double* __restrict__ Hm, float* __restrict__ Wm, size_t ldw,
                                                      size_t Np) {
  __shared__ double hs[kBMaxBasis * B];
  for (int t = threadIdx.x; t < n * B; t += blockDim.x) hs[t] = Hm[t];
  __syncthreads();
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    double acc[B];
#pragma unroll
    for (int r = 0; r < B; r++) acc[r] = Wm[size_t(r) * ldw + p];
    for (int i = 0; i < n; i++) {
      const double v = V[size_t(i) * ld + p];
#pragma unroll
      for (int r = 0; r < B; r++) acc[r] -= hs[i * B + r] * v;
    }
#pragma unroll
    for (int r = 0; r < B; r++) Wm[size_t(r) * ldw + p] = float(acc[r]);
  }
}

// In place X_r <- sum_s X_s Rinv[s][r] (upper triangular Rinv), FP64.
template <int B>
__global__ __launch_bounds__(256) void btransform_kernel(float* __restrict__ X, size_t ld, const double* __restrict__ Rinv,
                                                         size_t Np) {
  __shared__ double rs[B * B];
  for (int t = threadIdx.x; t < B * B; t += blockDim.x) rs[t] = Rinv[t];
  __syncthreads();
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    double x[B];
#pragma unroll
    for (int r = 0; r < B; r++) x[r] = X[size_t(r) * ld + p];
#pragma unroll
    for (int r = 0; r < B; r++) {
      double a = 0;
#pragma unroll
      for (int q = 0; q <= r; q++) a += x[q] * rs[q * B + r];
      X[size_t(r) * ld + p] = float(a);
    }
  }
}

// Cholesky G = R^T R (upper R) of the B x B Gram matrix, Rinv = R^-1.
// pass 1 stores R into R1; pass 2 stores R2 * R1 into Rtot.  Pivots below
// 1e-12 of the largest diagonal entry are floored and counted in *brk.
template <int B>
__global__ void bchol_kernel(const double* __restrict__ G, double* __restrict__ Rinv, double* __restrict__ R1,
                             double* __restrict__ Rtot, int pass, int* __restrict__ brk) {
  if (threadIdx.x != 0) return;
  double R[B * B], Ri[B * B];
  double gmax = 0;
  for (int r = 0; r < B; r++) gmax = fmax(gmax, G[r * B + r]);
  for (int t = 0; t < B * B; t++) R[t] = Ri[t] = 0;
  for (int j = 0; j < B; j++) {
    double d = G[j * B + j];
    for (int k = 0; k < j; k++) d -= R[k * B + j] * R[k * B + j];
    if (!(d > 1e-12 * gmax)) {
      d = fmax(1e-12 * gmax, 1e-300);
      atomicAdd(brk, 1);
    }
    const double rjj = sqrt(d);
    R[j * B + j] = rjj;
    for (int i = j + 1; i < B; i++) {
      double a = 0.5 * (G[j * B + i] + G[i * B + j]);
      for (int k = 0; k < j; k++) a -= R[k * B + j] * R[k * B + i];
      R[j * B + i] = a / rjj;
    }
  }
  // inverse of the upper triangular R
  for (int j = 0; j < B; j++) {
    Ri[j * B + j] = 1.0 / R[j * B + j];
    for (int i = j - 1; i >= 0; i--) {
      double a = 0;
      for (int k = i + 1; k <= j; k++) a += R[i * B + k] * Ri[k * B + j];
      Ri[i * B + j] = -a / R[i * B + i];
    }
  }
  for (int t = 0; t < B * B; t++) Rinv[t] = Ri[t];
  if (pass == 1) {
    for (int t = 0; t < B * B; t++) R1[t] = R[t];
  } else {
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        double a = 0;
        for (int k = 0; k < B; k++) a += R[i * B + k] * R1[k * B + j];
        Rtot[i * B + j] = a;
      }
  }
}

// T[i][k + r] += H1[i][r] + H2[i][r] for i < n, then mirrored: rows above the
// panel copy their column entries, and the panel's diagonal block is made
// symmetric by averaging.  One block; the phases are separated by barriers so
// that no entry is read while another thread updates it.
__global__ void btcol_kernel(double* __restrict__ T, int M, int k, int n, int B, const double* __restrict__ H1,
                             const double* __restrict__ H2) {
  for (int t = threadIdx.x; t < n * B; t += blockDim.x) {
    const int i = t / B, r = t % B;
    T[size_t(i) * M + k + r] += H1[t] + H2[t];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < k * B; t += blockDim.x) {
    const int i = t / B, r = t % B;
    T[size_t(k + r) * M + i] = T[size_t(i) * M + k + r];
  }
  for (int t = threadIdx.x; t < B * B; t += blockDim.x) {
    const int a = t / B, r = t % B;
    if (a < r) {
      const double v = 0.5 * (T[size_t(k + a) * M + k + r] + T[size_t(k + r) * M + k + a]);
      T[size_t(k + a) * M + k + r] = v;
      T[size_t(k + r) * M + k + a] = v;
    }
  }
}

__global__ void fill_f64_kernel(double* p, size_t n, double v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

__global__ void fill_i32_kernel(int* p, int v) { *p = v; }

__global__ void t00_kernel(double* T, const double* h) { T[0] = h[0]; }

// Copies a caller-supplied block (generalised vectors v, padded) into
// u-space: u = D^1/2 v.
__global__ __launch_bounds__(256) void to_uspace_kernel(const float* __restrict__ v, const float* __restrict__ D,
                                                        float* __restrict__ u, size_t Np) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    u[p] = v[p] * sqrtf(D[p]);
}

namespace {

template <int B>
struct BlockOps {
  Context& c;
  const Stencil& st;
  size_t Np;
  int NC;
  double* partial;
  BSpmvArgs A;
  unsigned vg;
  BlockOps(Context& cc, const Stencil& s_) : c(cc), st(s_), Np(s_.Np) {
    NC = int((Np + kBDotChunk - 1) / kBDotChunk);
    partial = c.ws<double>("beig.partial", size_t(kBMaxBasis + kBMax) * kBMax * NC);
    int dx[kStencilHalf], dy[kStencilHalf];
    stencil_upper_offsets(dx, dy);
    for (int j = 0; j < kStencilHalf; j++) {
      A.o[j] = st.off_host[j];
      A.lo[j] = dy[j] * kBTileW + dx[j];
    }
    vg = std::min<unsigned>(2048, div_up(Np, 256));
  }
  void spmv(const float* X, float* Y) {
    KScope ks(c, kKSpmv, double(st.W) * st.H * (4.0 * kStencilHalf + 4 + 8.0 * B));
    hipLaunchKernelGGL(bspmv_kernel<B>, dim3(div_up(st.W, kBTX), div_up(st.H, kBTY)), dim3(kBTX, kBTY), 0, c.stream,
                       st.wup, st.s, X, Np, Y, Np, st.W, st.H, st.Wp, st.Hp, Np, A);
    GPB_LAUNCH_CHECK();
  }
  // dst[i * B + r] = V_i . (D) W_r, i < n
  void dots(const float* V, int n, const float* Wm, const float* D, double* dst) {
    {
      KScope ks(c, kKDots, double(Np) * (4.0 * n + 4.0 * B + (D ? 4 : 0)));
      hipLaunchKernelGGL(bdots_kernel<B>, dim3(NC, n), dim3(256), 0, c.stream, V, Np, Wm, Np, D, Np, partial);
      GPB_LAUNCH_CHECK();
    }
    launch_reduce(c, partial, NC, n * B, dst);
  }
  void update(const float* V, int n, const double* Hm, float* Wm) {
    KScope ks(c, kKUpdate, double(Np) * (4.0 * n + 8.0 * B));
    hipLaunchKernelGGL(bupdate_kernel<B>, dim3(vg), dim3(256), 0, c.stream, V, Np, n, Hm, Wm, Np, Np);
    GPB_LAUNCH_CHECK();
  }
  // CholQR2 of the panel X (in place); Rtot = R of X_in = X_out Rtot.
  void cholqr2(float* X, double* G, double* Rinv, double* R1, double* Rtot, int* brk) {
    for (int pass = 1; pass <= 2; pass++) {
      dots(X, B, X, nullptr, G);
      hipLaunchKernelGGL(bchol_kernel<B>, dim3(1), dim3(64), 0, c.stream, G, Rinv, R1, Rtot, pass, brk);
      GPB_LAUNCH_CHECK();
      KScope ks(c, kKUpdate, double(Np) * 8.0 * B);
      hipLaunchKernelGGL(btransform_kernel<B>, dim3(vg), dim3(256), 0, c.stream, X, Np, Rinv, Np);
      GPB_LAUNCH_CHECK();
    }
  }
};

template <int B>
void block_solve_impl(Context& c, const Stencil& st, const EigOptions& opt_in, const BlockOptions& bo, double* evals,
                      float* d_evecs, LanczosStats& stats) {
  EigOptions opt = opt_in;
  const size_t Np = st.Np;
  const size_t N = size_t(st.W) * st.H;
  const int K = opt.K;
  if (K < 2) throw ArgError("need at least 2 eigenpairs");
  // M: projected size (vectors with complete columns); capacity M + B
  int M = std::min(opt.m, kBMaxBasis - B);
  M = std::min<int>(M, int(N) - B - 1);
  int keep = std::min(std::max(opt.keep, K + 1), M - B);
  if (M < K + 2 * B || keep < K) throw ArgError("basis too small for the block size");
  BlockOps<B> ops(c, st);
  float* V = c.ws<float>("beig.V", size_t(M + B) * Np);
  float* V2 = c.ws<float>("beig.V2", size_t(M + B) * Np);
  float* Wb = c.ws<float>("beig.W", size_t(B) * Np);
  double* T = c.ws<double>("beig.T", size_t(M) * M);
  double* H1 = c.ws<double>("beig.H1", size_t(kBMaxBasis + kBMax) * kBMax);
  double* H2 = c.ws<double>("beig.H2", size_t(kBMaxBasis + kBMax) * kBMax);
  double* G = c.ws<double>("beig.G", kBMax * kBMax);
  double* Rinv = c.ws<double>("beig.Rinv", kBMax * kBMax);
  double* R1 = c.ws<double>("beig.R1", kBMax * kBMax);
  double* Rtot = c.ws<double>("beig.R", kBMax * kBMax);
  double* Ydev = c.ws<double>("beig.Y", size_t(M) * M);
  double* nrm = c.ws<double>("beig.nrm", kBMaxBasis);
  int* brk = c.ws<int>("beig.brk", 1);
  hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(1), 0, c.stream, brk, 0);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(fill_f64_kernel, dim3(64), dim3(256), 0, c.stream, T, size_t(M) * M, 0.0);
  GPB_LAUNCH_CHECK();
  launch_fill(c, Wb, size_t(B) * Np, 0.f);

  // locked u0 = D^1/2 1 / ||.||, starting panel at [1, 1 + B)
  launch_locked_start(c, st, V, V + Np, B, Np, 0x5eed1234abcdull);
  ops.dots(V, 1, V, nullptr, nrm);  // uses B columns of V: [0, B); only entry 0 is used
  launch_scale_by(c, V, nrm, Np);
  if (bo.init_block) {
    for (int r = 0; r < B; r++) {
      hipLaunchKernelGGL(to_uspace_kernel, dim3(ops.vg), dim3(256), 0, c.stream, bo.init_block + size_t(r) * Np, st.D,
                         V + size_t(1 + r) * Np, Np);
      GPB_LAUNCH_CHECK();
    }
  }
  // T00 = u0' L u0
  ops.spmv(V, Wb);
  ops.dots(V, 1, Wb, nullptr, H1);
  hipLaunchKernelGGL(t00_kernel, dim3(1), dim3(1), 0, c.stream, T, H1);
  GPB_LAUNCH_CHECK();
  // orthonormalise the starting panel against u0 (twice), then CholQR2
  float* P0 = V + Np;
  for (int pass = 0; pass < 2; pass++) {
    ops.dots(V, 1, P0, nullptr, H1);
    ops.update(V, 1, H1, P0);
  }
  ops.cholqr2(P0, G, Rinv, R1, Rtot, brk);

  int k = 1;  // first column of the current panel
  stats = LanczosStats();
  stats.basis = M;
  stats.block = B;
  int matvecs = 1;
  std::vector<double> Th(size_t(M) * M), th, Y, Rh(B * B);
  std::vector<double> resid(K, 0.0);
  bool converged = false;
  int kcur = 0;
  while (true) {
    while (k + B <= M) {
      float* Pk = V + size_t(k) * Np;
      const int n = k + B;
      ops.spmv(Pk, Wb);
      matvecs += B;
      ops.dots(V, n, Wb, nullptr, H1);
      ops.update(V, n, H1, Wb);
      ops.dots(V, n, Wb, nullptr, H2);
      ops.update(V, n, H2, Wb);
      hipLaunchKernelGGL(btcol_kernel, dim3(1), dim3(256), 0, c.stream, T, M, k, n, B, H1, H2);
      GPB_LAUNCH_CHECK();
      ops.cholqr2(Wb, G, Rinv, R1, Rtot, brk);
      GPB_HIP(hipMemcpyAsync(V + size_t(n) * Np, Wb, size_t(B) * Np * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
      k += B;
    }
    kcur = k;  // projected size
    // Rayleigh-Ritz on the host (smallest first)
    GPB_HIP(hipMemcpyAsync(Th.data(), T, sizeof(double) * M * M, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipMemcpyAsync(Rh.data(), Rtot, sizeof(double) * B * B, hipMemcpyDeviceToHost, c.stream));
    // weighted Gram of the residual panel
    ops.dots(V + size_t(kcur) * Np, B, V + size_t(kcur) * Np, st.D, G);
    std::vector<double> Gd(B * B);
    int brk_h = 0;
    GPB_HIP(hipMemcpyAsync(Gd.data(), G, sizeof(double) * B * B, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipMemcpyAsync(&brk_h, brk, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    c.kt.fold();
    // A rank-deficient panel means the block Krylov space became (nearly)
    // invariant; the caller falls back to the single-vector solver.
    if (brk_h > 0) {
      stats.iterations = matvecs;
      throw BlockBreakdown("block Lanczos: rank-deficient panel");
    }
    std::vector<double> A(size_t(kcur) * kcur);
    for (int i = 0; i < kcur; i++)
      for (int j = 0; j < kcur; j++) A[size_t(i) * kcur + j] = Th[size_t(i) * M + j];
    {
      std::vector<double> thd, Yd;
      symeig_desc(kcur, A, thd, Yd);
      th.assign(size_t(kcur), 0.0);
      Y.assign(size_t(kcur) * kcur, 0.0);
      for (int i = 0; i < kcur; i++) {
        th[size_t(i)] = thd[size_t(kcur - 1 - i)];
        for (int j = 0; j < kcur; j++) Y[size_t(j) * kcur + i] = Yd[size_t(j) * kcur + (kcur - 1 - i)];
      }
    }
    const int nk = keep;
    std::vector<double> Yk(size_t(kcur) * nk);
    for (int j = 0; j < kcur; j++)
      for (int i = 0; i < nk; i++) Yk[size_t(j) * nk + i] = Y[size_t(j) * kcur + i];
    GPB_HIP(hipMemcpyAsync(Ydev, Yk.data(), sizeof(double) * kcur * nk, hipMemcpyHostToDevice, c.stream));
    launch_ritz(c, V, Np, kcur, nk, Ydev, V2, Np);
    GPB_HIP(hipMemcpyAsync(V2 + size_t(nk) * Np, V + size_t(kcur) * Np, size_t(B) * Np * sizeof(float),
                           hipMemcpyDeviceToDevice, c.stream));
    // ||D^1/2 u_i|| for the wanted Ritz vectors
    for (int i = 0; i < K; i++) ops.dots(V2 + size_t(i) * Np, 1, V2 + size_t(i) * Np, st.D, nrm + i * B);
    std::vector<double> nh(size_t(K) * B);
    GPB_HIP(hipMemcpyAsync(nh.data(), nrm, sizeof(double) * K * B, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    std::swap(V, V2);
    double worst = 0;
    for (int i = 0; i < K; i++) {
      // residual r_i = V_res (R Y_last,i)
      double cvec[kBMax];
      for (int a = 0; a < B; a++) {
        double s = 0;
        for (int b2 = 0; b2 < B; b2++) s += Rh[size_t(a) * B + b2] * Y[size_t(kcur - B + b2) * kcur + i];
        cvec[a] = s;
      }
      double rr = 0;
      for (int a = 0; a < B; a++)
        for (int b2 = 0; b2 < B; b2++) rr += cvec[a] * Gd[size_t(a) * B + b2] * cvec[b2];
      const double un = nh[size_t(i) * B];
      resid[size_t(i)] = std::sqrt(std::max(rr, 0.0)) / std::sqrt(std::max(un, 1e-300));
      worst = std::max(worst, resid[size_t(i)]);
    }
    stats.max_residual = worst;
    converged = worst <= opt.tol;
    if (converged || matvecs >= opt.max_matvecs) break;
    std::vector<double> T0(size_t(M) * M, 0.0);
    for (int i = 0; i < keep; i++) T0[size_t(i) * M + i] = th[size_t(i)];
    GPB_HIP(hipMemcpyAsync(T, T0.data(), sizeof(double) * M * M, hipMemcpyHostToDevice, c.stream));
    k = keep;
    stats.restarts++;
  }
  stats.iterations = matvecs;
  stats.converged = 0;
  for (int i = 0; i < K; i++) stats.converged += resid[size_t(i)] <= opt.tol;
  for (int i = 0; i < K; i++) {
    evals[i] = th[size_t(i)];
    launch_to_general(c, V + size_t(i) * Np, st.s, d_evecs + size_t(i) * Np, Np);
  }
  if (!converged)
    throw NoConvergence("block Lanczos stopped after " + std::to_string(matvecs) +
                        " operator applications with residual " + std::to_string(stats.max_residual));
}

}  // namespace

// Diagnostic: Y = L X for B columns with the block kernel, or the
// single-vector kernel column by column when B == 1.
void debug_block_spmv(Context& c, const Stencil& st, int B, const float* X, float* Y) {
  switch (B) {
    case 1:
      launch_spmv(c, st, X, Y);
      break;
    case 4: {
      BlockOps<4> ops(c, st);
      ops.spmv(X, Y);
    } break;
    case 8: {
      BlockOps<8> ops(c, st);
      ops.spmv(X, Y);
    } break;
    case 16: {
      BlockOps<16> ops(c, st);
      ops.spmv(X, Y);
    } break;
    default:
      throw ArgError("block size must be 1, 4, 8 or 16");
  }
}

void block_lanczos_solve(Context& c, const Stencil& st, const EigOptions& opt, const BlockOptions& bo, double* evals,
                         float* d_evecs, LanczosStats& stats) {
  switch (bo.block) {
    case 4:
      block_solve_impl<4>(c, st, opt, bo, evals, d_evecs, stats);
      break;
    case 8:
      block_solve_impl<8>(c, st, opt, bo, evals, d_evecs, stats);
      break;
    case 12:
      block_solve_impl<12>(c, st, opt, bo, evals, d_evecs, stats);
      break;
    case 16:
      block_solve_impl<16>(c, st, opt, bo, evals, d_evecs, stats);
      break;
    default:
      throw ArgError("block size must be 4, 8, 12 or 16");
  }
}

}  // namespace gpb

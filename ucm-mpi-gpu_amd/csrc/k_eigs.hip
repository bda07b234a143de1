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
#include <cmath>
#include <cstring>
#include <vector>

#include "gpb_pipeline.hpp"

namespace gpb {

void symeig_desc(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& z);

constexpr int kRedThreads = 256;
constexpr int kMaxBasis = 128;

__global__ void fill_f32_kernel(float* __restrict__ p, size_t n, float v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

void launch_fill(Context& c, float* p, size_t n, float v) {
  if (!n) return;
  hipLaunchKernelGGL(fill_f32_kernel, dim3(std::min<unsigned>(4096, div_up(n, 256))), dim3(256), 0, c.stream, p, n, v);
  GPB_LAUNCH_CHECK();
}

struct SpmvOffsets {
  int o[kStencilHalf];
};

// Normalised Laplacian in difference form: with u = s .* x (s = D^-1/2),
// y_p = s_p * sum_q W_pq (u_p - u_q).  Differences of neighbouring values are
// formed before the products, so eigenvalues near 0 keep their relative
// accuracy in FP32 (the form I - D^-1/2 W D^-1/2 would lose it to cancellation
// against 1).  W is read once per unordered pair: W[k][p] for q = p + e_k and
// W[k][p - e_k] for q = p - e_k.
__global__ __launch_bounds__(256) void spmv_kernel(const float* __restrict__ wup, const float* __restrict__ s,
                                                   const float* __restrict__ x, float* __restrict__ y, int W, int H,
                                                   int Wp, size_t Np, SpmvOffsets A) {
  const int xi = blockIdx.x * blockDim.x + threadIdx.x, yi = blockIdx.y * blockDim.y + threadIdx.y;
  if (xi >= W || yi >= H) return;
  const size_t p = size_t(yi + kICRadius) * Wp + xi + kICRadius;
  const float sp = s[p];
  const float up = sp * x[p];
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < kStencilHalf; j++) {
    const int o = A.o[j];
    acc = __builtin_fmaf(wup[size_t(j) * Np + p], up - s[p + o] * x[p + o], acc);
    acc = __builtin_fmaf(wup[size_t(j) * Np + p - o], up - s[p - o] * x[p - o], acc);
  }
  y[p] = sp * acc;
}

void launch_spmv(Context& c, const Stencil& st, const float* x, float* y) {
  SpmvOffsets A;
  for (int j = 0; j < kStencilHalf; j++) A.o[j] = st.off_host[j];
  dim3 block(64, 4), grid(div_up(st.W, 64), div_up(st.H, 4));
  KScope ks(c, kKSpmv, double(st.W) * st.H * (4 * kStencilHalf + 12));
  hipLaunchKernelGGL(spmv_kernel, grid, block, 0, c.stream, st.wup, st.s, x, y, st.W, st.H, st.Wp, st.Np, A);
  GPB_LAUNCH_CHECK();
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < int(blockDim.x >> 6); i++) r += sh[i];
  return r;
}

// partial[i][blk] = sum over the block's chunk of V_i . w (FP64), i < n;
// V rows are ld apart.  If wsq, partial[n][blk] = w . w (with weights if D).
__global__ __launch_bounds__(kRedThreads) void multidot_kernel(const float* __restrict__ V, size_t ld, int n,
                                                               const float* __restrict__ w, const float* __restrict__ D,
                                                               size_t Np, int wsq, double* __restrict__ partial) {
  __shared__ double sh[kRedThreads / 64];
  const size_t per = (Np + gridDim.x - 1) / gridDim.x;
  const size_t lo = size_t(blockIdx.x) * per, hi = lo + per < Np ? lo + per : Np;
  for (int i0 = 0; i0 < n; i0 += 4) {
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    const float* v0 = V + size_t(i0) * ld;
    for (size_t p = lo + threadIdx.x; p < hi; p += kRedThreads) {
      const double wv = w[p];
      a0 += double(v0[p]) * wv;
      if (i0 + 1 < n) a1 += double(v0[ld + p]) * wv;
      if (i0 + 2 < n) a2 += double(v0[2 * ld + p]) * wv;
      if (i0 + 3 < n) a3 += double(v0[3 * ld + p]) * wv;
    }
    double r;
    r = block_sum(a0, sh);
    if (threadIdx.x == 0) partial[size_t(i0) * gridDim.x + blockIdx.x] = r;
    if (i0 + 1 < n) {
      r = block_sum(a1, sh);
      if (threadIdx.x == 0) partial[size_t(i0 + 1) * gridDim.x + blockIdx.x] = r;
    }
    if (i0 + 2 < n) {
      r = block_sum(a2, sh);
      if (threadIdx.x == 0) partial[size_t(i0 + 2) * gridDim.x + blockIdx.x] = r;
    }
    if (i0 + 3 < n) {
      r = block_sum(a3, sh);
      if (threadIdx.x == 0) partial[size_t(i0 + 3) * gridDim.x + blockIdx.x] = r;
    }
  }
  if (wsq) {
    double a = 0;
    for (size_t p = lo + threadIdx.x; p < hi; p += kRedThreads) {
      const double wv = w[p];
      a += (D ? double(D[p]) : 1.0) * wv * wv;
    }
    const double r = block_sum(a, sh);
    if (threadIdx.x == 0) partial[size_t(n) * gridDim.x + blockIdx.x] = r;
  }
}

// out[i] = sum_blk partial[i][blk]  (one block per i, fixed order)
__global__ __launch_bounds__(kRedThreads) void reduce_kernel(const double* __restrict__ partial, int G,
                                                             double* __restrict__ out) {
  __shared__ double sh[kRedThreads / 64];
  double a = 0;
  for (int b = threadIdx.x; b < G; b += kRedThreads) a += partial[size_t(blockIdx.x) * G + b];
  const double r = block_sum(a, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// w <- w - sum_i h_i V_i  (FP64 accumulation), over the whole padded range
__global__ __launch_bounds__(256) void update_kernel(const float* __restrict__ V, size_t ld, int n,
                                                     const double* __restrict__ h, float* __restrict__ w, size_t Np) {
  __shared__ double hs[kMaxBasis];
  for (int i = threadIdx.x; i < n; i += blockDim.x) hs[i] = h[i];
  __syncthreads();
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    double acc = w[p];
    for (int i = 0; i < n; i++) acc -= hs[i] * double(V[size_t(i) * ld + p]);
    w[p] = float(acc);
  }
}

// T[i][j] += h[i] (and T[j][i]) for i <= j; on the final pass also beta.
__global__ void tcol_kernel(double* __restrict__ T, int m, int j, const double* __restrict__ h, int final_pass,
                            const double* __restrict__ nrm, double* __restrict__ beta) {
  const int i = threadIdx.x;
  if (i <= j) {
    T[size_t(i) * m + j] += h[i];
    if (i != j) T[size_t(j) * m + i] = T[size_t(i) * m + j];
  }
  if (final_pass && i == 0) {
    double b = sqrt(nrm[0]);
    beta[j] = b;
    beta[kMaxBasis + j] = b > 1e-300 ? 1.0 / b : 0.0;
  }
}

__global__ __launch_bounds__(256) void scale_kernel(const float* __restrict__ w, const double* __restrict__ beta, int j,
                                                    float* __restrict__ v, size_t Np) {
  const double ib = beta[kMaxBasis + j];
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    v[p] = float(double(w[p]) * ib);
}

// out[i] = sum_j Y[j][i] V[j], i < nk, j < m (FP64 accumulation)
__global__ __launch_bounds__(256) void ritz_kernel(const float* __restrict__ V, size_t ld, int m, int nk,
                                                   const double* __restrict__ Y /* m x nk */, float* __restrict__ out,
                                                   size_t Np) {
  __shared__ double ys[kMaxBasis * 32];
  for (int i0 = 0; i0 < nk; i0 += 32) {
    const int cnt = nk - i0 < 32 ? nk - i0 : 32;
    __syncthreads();
    for (int t = threadIdx.x; t < m * cnt; t += blockDim.x) ys[t] = Y[size_t(t / cnt) * nk + i0 + (t % cnt)];
    __syncthreads();
    for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
      double acc[32];
#pragma unroll
      for (int i = 0; i < 32; i++) acc[i] = 0;
      for (int j = 0; j < m; j++) {
        const double v = V[size_t(j) * ld + p];
#pragma unroll
        for (int i = 0; i < 32; i++)
          if (i < cnt) acc[i] += ys[j * cnt + i] * v;
      }
#pragma unroll
      for (int i = 0; i < 32; i++)
        if (i < cnt) out[size_t(i0 + i) * ld + p] = float(acc[i]);
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

namespace {

struct Reducer {
  Context& c;
  size_t Np;
  int G;
  double* partial;
  double* out;
  Reducer(Context& cc, size_t np) : c(cc), Np(np) {
    G = int(std::min<size_t>(1024, std::max<size_t>(1, (np + 4095) / 4096)));
    partial = c.ws<double>("eig.partial", size_t(kMaxBasis + 2) * G);
    out = c.ws<double>("eig.redout", kMaxBasis + 2);
  }
  // out[0..n) = V_i . w, out[n] = w . (D) w if wsq
  void dots(const float* V, size_t ld, int n, const float* w, const float* D, bool wsq, double* dst) {
    KScope ks(c, kKDots, double(Np) * (4.0 * n + 4 + (wsq && D ? 4 : 0)));
    hipLaunchKernelGGL(multidot_kernel, dim3(G), dim3(kRedThreads), 0, c.stream, V, ld, n, w, D, Np, int(wsq), partial);
    GPB_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_kernel, dim3(n + (wsq ? 1 : 0)), dim3(kRedThreads), 0, c.stream, partial, G, dst);
    GPB_LAUNCH_CHECK();
  }
};

unsigned vec_grid(size_t Np) { return std::min<unsigned>(2048, div_up(Np, 256)); }

}  // namespace

void lanczos_solve(Context& c, const Stencil& st, const EigOptions& opt_in, double* evals, float* d_evecs,
                   LanczosStats& stats) {
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
  double* beta = c.ws<double>("eig.beta", 2 * kMaxBasis);
  double* Ydev = c.ws<double>("eig.Y", size_t(m) * m);
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
  hipLaunchKernelGGL(tcol_kernel, dim3(1), dim3(64), 0, c.stream, T, m, 0, h, 0, h, beta);
  GPB_LAUNCH_CHECK();
  // V1: random, orthogonal to V0, normalised
  float* V1 = V + Np;
  hipLaunchKernelGGL(random_start_kernel, dim3(vg), dim3(256), 0, c.stream, V1, st.W, st.H, st.Wp, Np,
                     uint64_t(0x5eed1234abcdull));
  GPB_LAUNCH_CHECK();
  for (int pass = 0; pass < 2; pass++) {
    red.dots(V, Np, 1, V1, nullptr, false, h);
    hipLaunchKernelGGL(update_kernel, dim3(vg), dim3(256), 0, c.stream, V, Np, 1, h, V1, Np);
    GPB_LAUNCH_CHECK();
  }
  red.dots(V, Np, 0, V1, nullptr, true, h);
  hipLaunchKernelGGL(scale_by_kernel, dim3(vg), dim3(256), 0, c.stream, V1, h, Np);
  GPB_LAUNCH_CHECK();

  int k = 1;
  stats = LanczosStats();
  stats.basis = m;
  stats.block = 1;
  int matvecs = 1;
  std::vector<double> Th(size_t(m) * m), th, Y;
  std::vector<double> betah(kMaxBasis);
  bool converged = false;
  std::vector<double> resid(K, 0.0);
  while (true) {
    for (int j = k; j < m; j++) {
      float* vj = V + size_t(j) * Np;
      launch_spmv(c, st, vj, w);
      matvecs++;
      for (int pass = 0; pass < 2; pass++) {
        const bool last = pass == 1;
        red.dots(V, Np, j + 1, w, nullptr, false, h);
        {
          KScope ks(c, kKUpdate, double(Np) * (4.0 * (j + 1) + 8));
          hipLaunchKernelGGL(update_kernel, dim3(vg), dim3(256), 0, c.stream, V, Np, j + 1, h, w, Np);
          GPB_LAUNCH_CHECK();
        }
        if (last) red.dots(V, Np, 0, w, nullptr, true, h + (kMaxBasis + 1));
        hipLaunchKernelGGL(tcol_kernel, dim3(1), dim3(kMaxBasis), 0, c.stream, T, m, j, h, int(last),
                           h + (kMaxBasis + 1), beta);
        GPB_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(scale_kernel, dim3(vg), dim3(256), 0, c.stream, w, beta, j, V + size_t(j + 1) * Np, Np);
      GPB_LAUNCH_CHECK();
    }
    // Rayleigh-Ritz on the host
    GPB_HIP(hipMemcpyAsync(Th.data(), T, sizeof(double) * m * m, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipMemcpyAsync(betah.data(), beta, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    std::vector<double> A = Th;
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
    std::vector<double> Yk(size_t(m) * nk);
    for (int j = 0; j < m; j++)
      for (int i = 0; i < nk; i++) Yk[size_t(j) * nk + i] = Y[size_t(j) * m + i];
    GPB_HIP(hipMemcpyAsync(Ydev, Yk.data(), sizeof(double) * m * nk, hipMemcpyHostToDevice, c.stream));
    {
      KScope ks(c, kKRitz, double(Np) * 4.0 * (m + nk));
      hipLaunchKernelGGL(ritz_kernel, dim3(vg), dim3(256), 0, c.stream, V, Np, m, nk, Ydev, V2, Np);
      GPB_LAUNCH_CHECK();
    }
    GPB_HIP(hipMemcpyAsync(V2 + size_t(nk) * Np, V + size_t(m) * Np, Np * sizeof(float), hipMemcpyDeviceToDevice,
                           c.stream));
    // weighted norms ||D^1/2 u_i||, ||D^1/2 v_m||
    std::vector<double> nrm(K + 1);
    for (int i = 0; i < K; i++) red.dots(V2, Np, 0, V2 + size_t(i) * Np, st.D, true, h + i);
    red.dots(V2, Np, 0, V2 + size_t(nk) * Np, st.D, true, h + K);
    GPB_HIP(hipMemcpyAsync(nrm.data(), h, sizeof(double) * (K + 1), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    c.kt.fold();
    std::swap(V, V2);
    const double bm = betah[m - 1];
    double worst = 0;
    for (int i = 0; i < K; i++) {
      resid[i] = std::fabs(bm * Y[size_t(m - 1) * m + i]) * std::sqrt(nrm[K]) / std::sqrt(std::max(nrm[i], 1e-300));
      worst = std::max(worst, resid[i]);
    }
    stats.max_residual = worst;
    converged = worst <= opt.tol;
    if (converged || matvecs >= opt.max_matvecs) break;
    // thick restart: basis = Ritz vectors [0, keep) and v_m; T = diag(theta)
    std::vector<double> T0(size_t(m) * m, 0.0);
    for (int i = 0; i < keep; i++) T0[size_t(i) * m + i] = th[size_t(i)];
    GPB_HIP(hipMemcpyAsync(T, T0.data(), sizeof(double) * m * m, hipMemcpyHostToDevice, c.stream));
    k = keep;
    stats.restarts++;
  }
  stats.iterations = matvecs;
  stats.converged = 0;
  for (int i = 0; i < K; i++) stats.converged += resid[i] <= opt.tol;
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

// The all-FP64 eigensolve: thick-restart Lanczos with every vector, the
// stencil, the SpMV and the reductions in FP64.  It is the precision
// reference for the FP32-storage solver (k_eigs.hip) on the same W, not a
// production path: simple kernels, classical Gram-Schmidt twice on every step.
//
// Reference semantics: the generalised eigenproblem of spectralPb.cpp:186-244
// (Spectra SymGEigsSolver, FP64 throughout, Cholesky mode), here on the
// normalised Laplacian L = D^-1/2 (D - W) D^-1/2 in difference form with the
// exact null vector D^1/2 1 locked, as the FP32 solver does.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "eig_common.hpp"

namespace gpb {

namespace {

constexpr int kMaxM64 = 64;

// y_p = s_p sum_q W_pq (u_p - u_q), u = s .* x, one thread per interior pixel.
__global__ __launch_bounds__(256) void spmv64_kernel(const double* __restrict__ wup, const double* __restrict__ s,
                                                     const double* __restrict__ x, double* __restrict__ y, int W,
                                                     int H, int Wp, size_t Np, SpmvArgs A) {
  const int xi = blockIdx.x * blockDim.x + threadIdx.x, yi = blockIdx.y;
  if (xi >= W) return;
  const size_t p = size_t(yi + kICRadius) * Wp + xi + kICRadius;
  const double up = s[p] * x[p];
  double acc = 0;
#pragma unroll 8
  for (int j = 0; j < kStencilHalf; j++) {
    const size_t qa = p + A.o[j], qb = p - A.o[j];
    acc = fma(wup[size_t(j) * Np + p], up - s[qa] * x[qa], acc);
    acc = fma(wup[size_t(j) * Np + qb], up - s[qb] * x[qb], acc);
  }
  y[p] = s[p] * acc;
}

// partial[i][blk] = V_i . w (i < n), partial[n][blk] = w . w (grid-stride).
template <int NV>
__global__ __launch_bounds__(256) void dots64_kernel(const double* __restrict__ V, size_t ld, int n,
                                                     const double* __restrict__ w, size_t Np,
                                                     double* __restrict__ partial) {
  __shared__ double sh[4][NV + 1];
  double acc[NV + 1];
#pragma unroll
  for (int i = 0; i <= NV; i++) acc[i] = 0;
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    const double wp = w[p];
    acc[NV] += wp * wp;
    double vv[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) vv[i] = V[size_t(i < n ? i : 0) * ld + p];
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] += vv[i] * wp;
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

__global__ __launch_bounds__(256) void reduce64_kernel(const double* __restrict__ partial, int G,
                                                       double* __restrict__ out) {
  __shared__ double red[4];
  double a = 0;
  for (int b = threadIdx.x; b < G; b += 256) a += partial[size_t(blockIdx.x) * G + b];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// w -= V h (i < n)
__global__ __launch_bounds__(256) void update64_kernel(const double* __restrict__ V, size_t ld, int n,
                                                       const double* __restrict__ h, double* __restrict__ w,
                                                       size_t Np) {
  __shared__ double hs[kMaxM64 + 1];
  for (int i = threadIdx.x; i < n; i += blockDim.x) hs[i] = h[i];
  __syncthreads();
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    double acc = w[p];
    for (int i = 0; i < n; i++) acc -= hs[i] * V[size_t(i) * ld + p];
    w[p] = acc;
  }
}

// out = w * f, f = 1 / sqrt(*nrm2) (or *inv when non-null)
__global__ __launch_bounds__(256) void scale64_kernel(const double* __restrict__ w, const double* __restrict__ nrm2,
                                                      double* __restrict__ out, size_t Np) {
  const double f = 1.0 / sqrt(*nrm2);
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    out[p] = w[p] * f;
}

// out_i = sum_j Y[j][i] V_j (i < nk <= 24)
__global__ __launch_bounds__(256) void ritz64_kernel(const double* __restrict__ V, size_t ld, int m, int nk,
                                                     const double* __restrict__ Y, double* __restrict__ out,
                                                     size_t Np) {
  constexpr int kC = 24;
  for (int i0 = 0; i0 < nk; i0 += kC) {
    const int cnt = nk - i0 < kC ? nk - i0 : kC;
    for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
      double acc[kC];
#pragma unroll
      for (int i = 0; i < kC; i++) acc[i] = 0;
      for (int j = 0; j < m; j++) {
        const double v = V[size_t(j) * ld + p];
        const double* y = Y + size_t(j) * nk + i0;
#pragma unroll
        for (int i = 0; i < kC; i++)
          if (i < cnt) acc[i] += y[i] * v;
      }
#pragma unroll
      for (int i = 0; i < kC; i++)
        if (i < cnt) out[size_t(i0 + i) * ld + p] = acc[i];
    }
  }
}

__global__ __launch_bounds__(256) void start64_kernel(const double* __restrict__ D, double* __restrict__ v0,
                                                      double* __restrict__ v1, int W, int H, int Wp, size_t Np,
                                                      uint64_t seed) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x) {
    v0[p] = sqrt(D[p]);
    const int y = int(p / Wp) - kICRadius, x = int(p % Wp) - kICRadius;
    double r = 0;
    if (x >= 0 && x < W && y >= 0 && y < H) {
      uint64_t z = seed + p * 0x9E3779B97F4A7C15ull;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      r = double(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    v1[p] = r;
  }
}

// v = s .* u (generalised eigenvectors), rows 0..K-1
__global__ __launch_bounds__(256) void general64_kernel(const double* __restrict__ V, const double* __restrict__ s,
                                                        int K, size_t Np, double* __restrict__ out) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
    for (int k = 0; k < K; k++) out[size_t(k) * Np + p] = V[size_t(k) * Np + p] * s[p];
}

// sum_p D_p V_i[p]^2 for rows rows[0..nr)
__global__ __launch_bounds__(256) void wnorm64_kernel(const double* __restrict__ V, size_t ld,
                                                      const int* __restrict__ rows, int nr,
                                                      const double* __restrict__ D, size_t Np,
                                                      double* __restrict__ partial) {
  __shared__ double red[4];
  for (int k = 0; k < nr; k++) {
    const double* v = V + size_t(rows[k]) * ld;
    double a = 0;
    for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < Np; p += size_t(gridDim.x) * blockDim.x)
      a += D[p] * v[p] * v[p];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) partial[size_t(k) * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
  }
}

struct Ops64 {
  Context& c;
  const Stencil64& st;
  size_t Np;
  int G = 1024;
  double* partial;
  SpmvArgs A{};
  Ops64(Context& cc, const Stencil64& s) : c(cc), st(s), Np(s.Np) {
    partial = c.ws<double>("e64.partial", size_t(kMaxM64 + 2) * G);
    for (int j = 0; j < kStencilHalf; j++) A.o[j] = st.off_host[j];
  }
  void spmv(const double* x, double* y) {
    KScope ks(c, kKSpmv, double(st.W) * st.H * (8.0 * kStencilHalf + 24));
    hipLaunchKernelGGL(spmv64_kernel, dim3(div_up(st.W, 256), st.H), dim3(256), 0, c.stream, st.wup, st.s, x, y, st.W,
                       st.H, st.Wp, Np, A);
    GPB_LAUNCH_CHECK();
  }
  // dst[0..n) = V_i . w, dst[n] = w . w
  void dots(const double* V, int n, const double* w, double* dst) {
    {
      KScope ks(c, kKDots, double(Np) * 8.0 * (n + 1));
      const dim3 g(G), b(256);
#define GPB_D64(NV)                                                                                    \
  case NV / 8:                                                                                         \
    hipLaunchKernelGGL(dots64_kernel<NV>, g, b, 0, c.stream, V, Np, n, w, Np, partial);               \
    break;
      switch ((n + 7) / 8) {
        case 0:
        GPB_D64(8)
        GPB_D64(16)
        GPB_D64(24)
        GPB_D64(32)
        GPB_D64(40)
        GPB_D64(48)
        GPB_D64(56)
        GPB_D64(64)
        default:
          throw ArgError("FP64 solver basis too large");
      }
#undef GPB_D64
      GPB_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(reduce64_kernel, dim3(n + 1), dim3(256), 0, c.stream, partial, G, dst);
    GPB_LAUNCH_CHECK();
  }
  void update(const double* V, int n, const double* h, double* w) {
    KScope ks(c, kKUpdate, double(Np) * 8.0 * (n + 2));
    hipLaunchKernelGGL(update64_kernel, dim3(2048), dim3(256), 0, c.stream, V, Np, n, h, w, Np);
    GPB_LAUNCH_CHECK();
  }
  void scale(const double* w, const double* nrm2, double* out) {
    hipLaunchKernelGGL(scale64_kernel, dim3(2048), dim3(256), 0, c.stream, w, nrm2, out, Np);
    GPB_LAUNCH_CHECK();
  }
};

}  // namespace

void lanczos_solve_f64(Context& c, const Stencil64& st, const EigOptions& opt_in, double* evals, double* d_evecs,
                       LanczosStats& stats) {
  EigOptions opt = opt_in;
  const size_t Np = st.Np;
  const size_t N = size_t(st.W) * st.H;
  if (opt.K < 2) throw ArgError("need at least 2 eigenpairs");
  if (N < size_t(opt.K) + 4) throw ArgError("image too small for the requested eigenpairs");
  opt.m = std::min<int>(std::min<int>(opt.m, kMaxM64 - 1), int(N) - 1);
  if (opt.m < opt.K + 2) throw ArgError("basis too small");
  opt.keep = std::max(opt.K + 1, std::min(opt.keep, opt.m - 2));
  const int m = opt.m, K = opt.K, keep = opt.keep;
  Ops64 ops(c, st);
  double* V = c.ws<double>("e64.V", size_t(m + 1) * Np);
  double* V2 = c.ws<double>("e64.V2", size_t(m + 1) * Np);
  double* w = c.ws<double>("e64.w", Np);
  double* h1 = c.ws<double>("e64.h1", kMaxM64 + 2);
  double* h2 = c.ws<double>("e64.h2", kMaxM64 + 2);
  double* Ydev = c.ws<double>("e64.Y", size_t(m) * keep);
  int* rows_dev = c.ws<int>("e64.rows", size_t(K + 1));
  launch_fill_f64(c, w, Np, 0.0);
  hipLaunchKernelGGL(start64_kernel, dim3(2048), dim3(256), 0, c.stream, st.D, V, V + Np, st.W, st.H, st.Wp, Np,
                     uint64_t(0x5eed1234abcdull));
  GPB_LAUNCH_CHECK();
  ops.dots(V, 1, V, h1);  // h1[1] = ||V0||^2
  ops.scale(V, h1 + 1, V);
  for (int pass = 0; pass < 2; pass++) {
    ops.dots(V, 1, V + Np, h1);
    ops.update(V, 1, h1, V + Np);
  }
  ops.dots(V, 1, V + Np, h1);
  ops.scale(V + Np, h1 + 1, V + Np);
  // T is assembled on the host from each step's projections (synchronous: the
  // reference path favours simplicity over overlap)
  std::vector<double> T(size_t(m) * m, 0.0), hh1(kMaxM64 + 2), hh2(kMaxM64 + 2);
  {
    ops.spmv(V, w);
    ops.dots(V, 1, w, h1);
    GPB_HIP(hipMemcpyAsync(hh1.data(), h1, sizeof(double) * 2, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    T[0] = hh1[0];
  }
  stats = LanczosStats();
  stats.basis = m;
  stats.block = 1;
  int matvecs = 1, k = 1;
  std::vector<double> th, Y, resid(size_t(K), 0.0);
  double beta_m = 0;
  std::vector<int> rows(size_t(K + 1));
  for (int i = 0; i < K; i++) rows[size_t(i)] = i;
  rows[size_t(K)] = keep;
  GPB_HIP(hipMemcpyAsync(rows_dev, rows.data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice, c.stream));
  while (true) {
    for (int j = k; j < m; j++) {
      const int n = j + 1;
      ops.spmv(V + size_t(j) * Np, w);
      matvecs++;
      ops.dots(V, n, w, h1);
      ops.update(V, n, h1, w);
      ops.dots(V, n, w, h2);
      ops.update(V, n, h2, w);
      ops.dots(V, 1, w, h2 + kMaxM64);  // h2[kMaxM64 + 1] = ||w||^2 after both passes
      ops.scale(w, h2 + kMaxM64 + 1, V + size_t(j + 1) * Np);
      GPB_HIP(hipMemcpyAsync(hh1.data(), h1, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
      GPB_HIP(hipMemcpyAsync(hh2.data(), h2, sizeof(double) * (kMaxM64 + 2), hipMemcpyDeviceToHost, c.stream));
      GPB_HIP(hipStreamSynchronize(c.stream));
      for (int i = 0; i <= j; i++) {
        const double v = T[size_t(i) * m + j] + hh1[size_t(i)] + hh2[size_t(i)];
        T[size_t(i) * m + j] = v;
        T[size_t(j) * m + i] = v;
      }
      beta_m = std::sqrt(hh2[size_t(kMaxM64 + 1)]);
    }
    std::vector<double> A = T, thd, Yd;
    symeig_desc(m, A, thd, Yd);
    th.assign(size_t(m), 0.0);
    Y.assign(size_t(m) * m, 0.0);
    for (int i = 0; i < m; i++) {
      th[size_t(i)] = thd[size_t(m - 1 - i)];
      for (int j = 0; j < m; j++) Y[size_t(j) * m + i] = Yd[size_t(j) * m + (m - 1 - i)];
    }
    std::vector<double> Yk(size_t(m) * keep);
    for (int j = 0; j < m; j++)
      for (int i = 0; i < keep; i++) Yk[size_t(j) * keep + i] = Y[size_t(j) * m + i];
    GPB_HIP(hipMemcpyAsync(Ydev, Yk.data(), sizeof(double) * m * keep, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(ritz64_kernel, dim3(2048), dim3(256), 0, c.stream, V, Np, m, keep, Ydev, V2, Np);
    GPB_LAUNCH_CHECK();
    GPB_HIP(hipMemcpyAsync(V2 + size_t(keep) * Np, V + size_t(m) * Np, Np * sizeof(double), hipMemcpyDeviceToDevice,
                           c.stream));
    hipLaunchKernelGGL(wnorm64_kernel, dim3(ops.G), dim3(256), 0, c.stream, V2, Np, rows_dev, K + 1, st.D, Np,
                       ops.partial);
    GPB_LAUNCH_CHECK();
    double* wn = c.ws<double>("e64.wn", size_t(K + 1));
    hipLaunchKernelGGL(reduce64_kernel, dim3(K + 1), dim3(256), 0, c.stream, ops.partial, ops.G, wn);
    GPB_LAUNCH_CHECK();
    std::vector<double> nrm(size_t(K + 1));
    GPB_HIP(hipMemcpyAsync(nrm.data(), wn, sizeof(double) * (K + 1), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    c.kt.fold();
    std::swap(V, V2);
    double worst = 0;
    for (int i = 0; i < K; i++) {
      resid[size_t(i)] = std::fabs(beta_m * Y[size_t(m - 1) * m + i]) * std::sqrt(nrm[size_t(K)]) /
                         std::sqrt(std::max(nrm[size_t(i)], 1e-300));
      worst = std::max(worst, resid[size_t(i)]);
    }
    stats.max_residual = worst;
    if (c.verbose)
      std::fprintf(stderr, "gpb: FP64 Lanczos %dx%d: %d operator applications, residual %.3e\n", st.W, st.H, matvecs,
                   worst);
    if (conv_count(opt, th.data(), resid.data()) == K || matvecs >= opt.max_matvecs) break;
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < keep; i++) T[size_t(i) * m + i] = th[size_t(i)];
    k = keep;
    stats.restarts++;
  }
  stats.iterations = matvecs;
  stats.converged = conv_count(opt, th.data(), resid.data());
  for (int i = 0; i < K; i++) evals[i] = th[size_t(i)];
  hipLaunchKernelGGL(general64_kernel, dim3(2048), dim3(256), 0, c.stream, V, st.s, K, Np, d_evecs);
  GPB_LAUNCH_CHECK();
  if (stats.converged < K)
    throw NoConvergence("FP64 Lanczos stopped after " + std::to_string(matvecs) +
                        " operator applications with residual " + std::to_string(stats.max_residual));
}

}  // namespace gpb

// FP64 Rayleigh-Ritz refinement of the eigenpairs a solver returns.
//
// The Lanczos solvers store their Krylov vectors in FP32: every stored
// w = L v carries a rounding of about 6e-8 of its entries, which puts noise of
// order 1e-10 (eps_32 / sqrt(N)) into the projected matrix and so into the
// Ritz values.  For the smallest eigenvalues of these operators (1e-5 and
// below) that is a relative error of a few 1e-5, whatever the residual
// target.  The refinement projects the pencil (D - W, D) onto the span of
// the K returned vectors with FP64 arithmetic throughout -- (D - W) v in the
// difference form sum_q W_pq (v_p - v_q), the degree 1 + sum_q W_pq summed in
// FP64 from the same stored W -- solves the K x K generalised problem and
// rotates the vectors.  Reference: the Ritz step of spectralPb.cpp:221-255 (Spectra),
// here restated as an explicit Rayleigh-Ritz in FP64.
#include <cmath>
#include <vector>

#include "gpb_pipeline.hpp"

namespace gpb {

void symeig_desc(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& z);

namespace {

constexpr int kRefineMax = 32;

struct RefineArgs {
  int o[kStencilHalf];
};

// Vector i = blockIdx.y: per workgroup partial sums over its pixels of
// g[j] = v_j . (D - W) v_i and m[j] = sum_p Dw_p v_i(p) v_j(p) for j < K,
// where Dw_p = 1 + sum_{q != p} W_pq in FP64 (the self-affinity is 1).
// partial[((i * 2K) + t) * G + block].
__global__ __launch_bounds__(256) void refine_gram_kernel(const float* __restrict__ wup, const float* __restrict__ V,
                                                          int K, int W, int H, int Wp, size_t Np, RefineArgs A,
                                                          double* __restrict__ partial) {
  __shared__ double sh[4][2 * kRefineMax];
  const int i = blockIdx.y;
  const float* vi = V + size_t(i) * Np;
  double acc[2 * kRefineMax];
#pragma unroll
  for (int t = 0; t < 2 * kRefineMax; t++) acc[t] = 0.0;
  const long long N = (long long)W * H;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < N; e += (long long)gridDim.x * 256) {
    const int y = int(e / W), x = int(e - (long long)y * W);
    const size_t p = size_t(y + kICRadius) * Wp + x + kICRadius;
    const double vp = vi[p];
    double yv = 0.0, d = 1.0;  // the degree counts the implicit self-affinity W_pp = 1
    for (int k = 0; k < kStencilHalf; k++) {
      const double w1 = wup[size_t(k) * Np + p], w2 = wup[size_t(k) * Np + p - A.o[k]];
      yv = fma(w1, vp - double(vi[p + A.o[k]]), yv);
      yv = fma(w2, vp - double(vi[p - A.o[k]]), yv);
      d += w1 + w2;
    }
#pragma unroll
    for (int j = 0; j < kRefineMax; j++) {
      if (j < K) {
        const double vj = V[size_t(j) * Np + p];
        acc[j] = fma(vj, yv, acc[j]);
        acc[kRefineMax + j] = fma(d * vp, vj, acc[kRefineMax + j]);
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < 2 * kRefineMax; t++) {
    if ((t < kRefineMax ? t : t - kRefineMax) < K) {
      double v = acc[t];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) sh[wv][t] = v;
    }
  }
  __syncthreads();
  const int G = gridDim.x;
  for (int t = threadIdx.x; t < 2 * K; t += 256) {
    const int src = t < K ? t : kRefineMax + (t - K);
    partial[(size_t(i) * 2 * K + t) * G + blockIdx.x] = sh[0][src] + sh[1][src] + sh[2][src] + sh[3][src];
  }
}

}  // namespace

void refine_pairs(Context& c, const Stencil& st, double* evals, float* d_evecs, int K) {
  if (K < 1 || K > kRefineMax) throw ArgError("refinement takes 1 to 32 pairs");
  const size_t Np = st.Np;
  RefineArgs A;
  for (int k = 0; k < kStencilHalf; k++) A.o[k] = st.off_host[k];
  const unsigned G = std::min<unsigned>(1024, div_up(size_t(st.W) * st.H, 256));
  double* part = c.ws<double>("refine.part", size_t(2) * K * K * G);
  double* sums = c.ws<double>("refine.sums", size_t(2) * K * K);
  {
    KScope ks(c, kKRitz, double(st.W) * st.H * K * (4.0 * (2 * kStencilHalf + 2 * kStencilHalf) + 4.0 * K));
    hipLaunchKernelGGL(refine_gram_kernel, dim3(G, K), dim3(256), 0, c.stream, st.wup, d_evecs, K, st.W, st.H, st.Wp,
                       Np, A, part);
    GPB_LAUNCH_CHECK();
  }
  launch_reduce(c, part, int(G), 2 * K * K, sums);
  std::vector<double> h(size_t(2) * K * K);
  GPB_HIP(hipMemcpyAsync(h.data(), sums, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
  // Gm[i][j] = v_j (D - W) v_i, Mm[i][j] = v_j D v_i, symmetrised
  std::vector<double> Gm(size_t(K) * K), Mm(size_t(K) * K);
  for (int i = 0; i < K; i++)
    for (int j = 0; j < K; j++) {
      Gm[size_t(i) * K + j] = 0.5 * (h[size_t(i) * 2 * K + j] + h[size_t(j) * 2 * K + i]);
      Mm[size_t(i) * K + j] = 0.5 * (h[size_t(i) * 2 * K + K + j] + h[size_t(j) * 2 * K + K + i]);
    }
  // Cholesky Mm = R^T R (upper R)
  std::vector<double> R(size_t(K) * K, 0.0);
  for (int j = 0; j < K; j++) {
    double d = Mm[size_t(j) * K + j];
    for (int k = 0; k < j; k++) d -= R[size_t(k) * K + j] * R[size_t(k) * K + j];
    if (!(d > 0)) return;  // the returned vectors are (numerically) dependent: keep the solver's pairs
    const double r = std::sqrt(d);
    R[size_t(j) * K + j] = r;
    for (int i = j + 1; i < K; i++) {
      double a = Mm[size_t(j) * K + i];
      for (int k = 0; k < j; k++) a -= R[size_t(k) * K + j] * R[size_t(k) * K + i];
      R[size_t(j) * K + i] = a / r;
    }
  }
  // Ri = R^-1 (upper); C = Ri^T Gm Ri
  std::vector<double> Ri(size_t(K) * K, 0.0);
  for (int j = 0; j < K; j++) {
    Ri[size_t(j) * K + j] = 1.0 / R[size_t(j) * K + j];
    for (int i = j - 1; i >= 0; i--) {
      double a = 0;
      for (int k = i + 1; k <= j; k++) a += R[size_t(i) * K + k] * Ri[size_t(k) * K + j];
      Ri[size_t(i) * K + j] = -a / R[size_t(i) * K + i];
    }
  }
  std::vector<double> T1(size_t(K) * K, 0.0), Cm(size_t(K) * K, 0.0);
  for (int i = 0; i < K; i++)
    for (int j = 0; j < K; j++) {
      double a = 0;
      for (int k = 0; k < K; k++) a += Gm[size_t(i) * K + k] * Ri[size_t(k) * K + j];
      T1[size_t(i) * K + j] = a;
    }
  for (int i = 0; i < K; i++)
    for (int j = 0; j < K; j++) {
      double a = 0;
      for (int k = 0; k < K; k++) a += Ri[size_t(k) * K + i] * T1[size_t(k) * K + j];
      Cm[size_t(i) * K + j] = a;
    }
  std::vector<double> th, Yd;
  symeig_desc(K, Cm, th, Yd);
  // ascending order; Z = Ri Y, the rotation of the vectors (column c: pair c)
  std::vector<double> Z(size_t(K) * K, 0.0);
  for (int cidx = 0; cidx < K; cidx++) {
    const int src = K - 1 - cidx;
    evals[cidx] = th[size_t(src)];
    for (int j = 0; j < K; j++) {
      double a = 0;
      for (int k = 0; k < K; k++) a += Ri[size_t(j) * K + k] * Yd[size_t(k) * K + src];
      Z[size_t(j) * K + cidx] = a;
    }
  }
  double* Zdev = c.ws<double>("refine.Z", size_t(K) * K);
  float* tmp = c.ws<float>("refine.V", size_t(K) * Np);
  GPB_HIP(hipMemcpyAsync(Zdev, Z.data(), sizeof(double) * Z.size(), hipMemcpyHostToDevice, c.stream));
  launch_ritz(c, d_evecs, Np, K, K, Zdev, tmp, Np);
  GPB_HIP(hipMemcpyAsync(d_evecs, tmp, sizeof(float) * K * Np, hipMemcpyDeviceToDevice, c.stream));
}

}  // namespace gpb

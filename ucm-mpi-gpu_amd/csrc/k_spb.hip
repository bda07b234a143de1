// Stage (e): eigenvector normalisation, sPb, gPb and the computeGPb output
// transforms.
//
// References: spectralPb.cpp:257-322 (min-max normalisation of eigenvectors
// 1..K-1, division by sqrt(lambda), 8 oriented derivative filters with
// BORDER_REFLECT_101, sPb_o += |response|), globalPb.cpp:31-37,97-103 (gPb
// weights), damascene globalPb.cu:5-49 (max over orientations masked by
// mPb > 0.05), skeleton.cu:229-249 (calibration sigmoid) and
// damascene.cu:594-604 (column-major oriented output).
#include "gpb_pipeline.hpp"

namespace gpb {

constexpr int kSpbT = 16;
constexpr int kSpbTile = kSpbT + 2 * kSPbRadius;  // 22

template <typename VT>
struct SpbParams {
  VT filt[kNumOri][kSPbSize * kSPbSize];
  const VT* evecs;         // [K][Np] padded
  const VT* mnmx;          // [K][2] min, max over the image
  VT inv_sqrt_lam[kMaxNvec];  // 0 => vector skipped
  int K, W, H, Wp;
  size_t Np;
  VT* spb;                 // [8][N]
};

template <typename VT>
__global__ __launch_bounds__(256) void minmax_kernel(const VT* __restrict__ v, int W, int H, int Wp,
                                                     VT* __restrict__ partial /* [2][G] */) {
  __shared__ VT smn[256], smx[256];
  VT mn = VT(__builtin_huge_val()), mx = -VT(__builtin_huge_val());
  const size_t N = size_t(W) * H;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += size_t(gridDim.x) * blockDim.x) {
    const int y = int(i / W), x = int(i % W);
    const VT a = v[size_t(y + kICRadius) * Wp + x + kICRadius];
    mn = a < mn ? a : mn;
    mx = a > mx ? a : mx;
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) {
      smn[threadIdx.x] = smn[threadIdx.x + s] < smn[threadIdx.x] ? smn[threadIdx.x + s] : smn[threadIdx.x];
      smx[threadIdx.x] = smx[threadIdx.x + s] > smx[threadIdx.x] ? smx[threadIdx.x + s] : smx[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = smn[0];
    partial[gridDim.x + blockIdx.x] = smx[0];
  }
}

template <typename VT>
__global__ __launch_bounds__(256) void minmax_final_kernel(const VT* __restrict__ partial, int G,
                                                           VT* __restrict__ out) {
  __shared__ VT smn[256], smx[256];
  VT mn = VT(__builtin_huge_val()), mx = -VT(__builtin_huge_val());
  for (int i = threadIdx.x; i < G; i += 256) {
    mn = partial[i] < mn ? partial[i] : mn;
    mx = partial[G + i] > mx ? partial[G + i] : mx;
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) {
      smn[threadIdx.x] = smn[threadIdx.x + s] < smn[threadIdx.x] ? smn[threadIdx.x + s] : smn[threadIdx.x];
      smx[threadIdx.x] = smx[threadIdx.x + s] > smx[threadIdx.x] ? smx[threadIdx.x + s] : smx[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smn[0];
    out[1] = smx[0];
  }
}

__device__ __forceinline__ int d_refl101(int i, int n) {
  if (n == 1) return 0;
  const int p = 2 * n - 2;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - i;
}

template <typename VT>
__global__ __launch_bounds__(kSpbT * kSpbT) void spb_kernel(SpbParams<VT> P) {
  __shared__ VT tile[kSpbTile * kSpbTile];
  const int tid = threadIdx.y * kSpbT + threadIdx.x;
  const int x0 = blockIdx.x * kSpbT - kSPbRadius, y0 = blockIdx.y * kSpbT - kSPbRadius;
  const int x = blockIdx.x * kSpbT + threadIdx.x, y = blockIdx.y * kSpbT + threadIdx.y;
  VT acc[kNumOri];
#pragma unroll
  for (int o = 0; o < kNumOri; o++) acc[o] = VT(0);
  for (int i = 1; i < P.K; i++) {
    const VT isl = P.inv_sqrt_lam[i];
    if (isl == VT(0)) continue;  // uniform across the grid
    const VT mn = P.mnmx[2 * i], mx = P.mnmx[2 * i + 1];
    const bool norm = mx > mn;
    const VT inv = norm ? VT(1) / (mx - mn) : VT(1);
    const VT* v = P.evecs + size_t(i) * P.Np;
    __syncthreads();
    for (int t = tid; t < kSpbTile * kSpbTile; t += kSpbT * kSpbT) {
      const int ty = t / kSpbTile, tx = t - ty * kSpbTile;
      const int gx = d_refl101(x0 + tx, P.W), gy = d_refl101(y0 + ty, P.H);
      VT a = v[size_t(gy + kICRadius) * P.Wp + gx + kICRadius];
      if (norm) a = (a - mn) * inv;
      tile[t] = a * isl;
    }
    __syncthreads();
    const VT* c = tile + threadIdx.y * kSpbTile + threadIdx.x;
#pragma unroll
    for (int o = 0; o < kNumOri; o++) {
      VT r = VT(0);
#pragma unroll
      for (int a = 0; a < kSPbSize; a++)
#pragma unroll
        for (int b = 0; b < kSPbSize; b++) r = fma_t(P.filt[o][a * kSPbSize + b], c[a * kSpbTile + b], r);
      acc[o] += r < VT(0) ? -r : r;
    }
  }
  if (x < P.W && y < P.H) {
    const size_t N = size_t(P.W) * P.H, p = size_t(y) * P.W + x;
#pragma unroll
    for (int o = 0; o < kNumOri; o++) P.spb[o * N + p] = acc[o];
  }
}

template <typename VT>
static void spb_t(Context& c, int W, int H, int Wp, size_t Np, const VT* d_evecs, const double* evals, int K,
                  VT* d_spb) {
  if (K > kMaxNvec) throw ArgError("too many eigenvectors");
  const int G = 256;
  VT* part = c.ws<VT>(sizeof(VT) == 8 ? "spb.part64" : "spb.part", 2 * G);
  VT* mnmx = c.ws<VT>(sizeof(VT) == 8 ? "spb.mnmx64" : "spb.mnmx", 2 * K);
  for (int i = 1; i < K; i++) {
    hipLaunchKernelGGL(minmax_kernel<VT>, dim3(G), dim3(256), 0, c.stream, d_evecs + size_t(i) * Np, W, H, Wp, part);
    GPB_LAUNCH_CHECK();
    hipLaunchKernelGGL(minmax_final_kernel<VT>, dim3(1), dim3(256), 0, c.stream, part, G, mnmx + 2 * i);
    GPB_LAUNCH_CHECK();
  }
  SpbParams<VT> P{};
  const std::vector<double>& f = spb_filters();
  for (int o = 0; o < kNumOri; o++)
    for (int t = 0; t < kSPbSize * kSPbSize; t++) P.filt[o][t] = VT(f[size_t(o) * kSPbSize * kSPbSize + t]);
  P.evecs = d_evecs;
  P.mnmx = mnmx;
  for (int i = 0; i < kMaxNvec; i++)
    P.inv_sqrt_lam[i] = (i > 0 && i < K && evals[i] > 2.220446049250313e-16) ? VT(1.0 / std::sqrt(evals[i])) : VT(0);
  P.K = K;
  P.W = W;
  P.H = H;
  P.Wp = Wp;
  P.Np = Np;
  P.spb = d_spb;
  KScope ks(c, kKSpb, double(W) * H * sizeof(VT) * (K - 1 + 8));
  hipLaunchKernelGGL(spb_kernel<VT>, dim3(div_up(W, kSpbT), div_up(H, kSpbT)), dim3(kSpbT, kSpbT), 0, c.stream, P);
  GPB_LAUNCH_CHECK();
}

void launch_spb(Context& c, const Stencil& st, const float* d_evecs, const double* evals, int K, float* d_spb) {
  spb_t<float>(c, st.W, st.H, st.Wp, st.Np, d_evecs, evals, K, d_spb);
}

void launch_spb_f64(Context& c, const Stencil64& st, const double* d_evecs, const double* evals, int K,
                    double* d_spb) {
  spb_t<double>(c, st.W, st.H, st.Wp, st.Np, d_evecs, evals, K, d_spb);
}

// gPb_o = gcue_o + w_sPb * sPb_o; max masked by mPb_nmax > 0.05; sigmoid maps
// written column-major through an LDS transpose, orientation o into block
// 7 - o: computeGPb's block i holds the slot oriMap[i] = (i + 4) mod 8
// (damascene/src/damascene.cu:598-603), and slot s holds the filter that
// damascene builds for angle pi/2 + (7 - s) pi/8 (spectralPb.cu:706-712) and
// applies row-major (spectralPb.cu:101), which is our orientation (3 - s) mod 8.
constexpr int kGT = 32;
__global__ __launch_bounds__(kGT * 8) void gpb_kernel(const float* __restrict__ gcue, const float* __restrict__ spb,
                                                      const float* __restrict__ nmax, int W, int H, float wspb,
                                                      float* __restrict__ orient, float* __restrict__ gmax,
                                                      float* __restrict__ gall) {
  __shared__ float tr[kNumOri][kGT][kGT + 1];
  const size_t N = size_t(W) * H;
  const int bx = blockIdx.x * kGT, by = blockIdx.y * kGT;
  for (int r = threadIdx.y; r < kGT; r += 8) {
    const int x = bx + threadIdx.x, y = by + r;
    if (x < W && y < H) {
      const size_t p = size_t(y) * W + x;
      float mv = -200.f;
#pragma unroll
      for (int o = 0; o < kNumOri; o++) {
        const float g = __builtin_fmaf(wspb, spb[o * N + p], gcue[o * N + p]);
        if (orient) orient[o * N + p] = g;
        mv = g > mv ? g : mv;
        float v = g;
        if (double(v) * 1.2 > 1) v = 1.f;
        else v = float(double(v) * 1.2);
        if (v < 0) v = 0.f;
        v = 1.0f / (1.0f + expf(float(2.6433 - 10.7998 * double(v))));
        v = float((double(v) - 0.0667) / 0.9333);
        tr[o][r][threadIdx.x] = v;
      }
      if (gmax) gmax[p] = nmax[p] > 0.05f ? mv : 0.f;
    }
  }
  if (!gall) return;
  __syncthreads();
  // column-major: gall[o][x * H + y]
  for (int r = threadIdx.y; r < kGT; r += 8) {
    const int x = bx + r, y = by + threadIdx.x;
    if (x < W && y < H)
#pragma unroll
      for (int o = 0; o < kNumOri; o++) gall[(kNumOri - 1 - o) * N + size_t(x) * H + y] = tr[o][threadIdx.x][r];
  }
}

void launch_gpb(Context& c, int W, int H, const float* d_gcue, const float* d_spb, const float* d_nmax,
                float* d_orient, float* d_gmax, float* d_gall) {
  KScope ks(c, kKGpb, double(W) * H * 104);
  hipLaunchKernelGGL(gpb_kernel, dim3(div_up(W, kGT), div_up(H, kGT)), dim3(kGT, 8), 0, c.stream, d_gcue, d_spb, d_nmax,
                     W, H, float(kGPbWeights[kNumCues]), d_orient, d_gmax, d_gall);
  GPB_LAUNCH_CHECK();
}

// The all-FP64 combine: orient[o] = gcue[o] + w_sPb * sPb[o]; gmax = max_o
// orient where mPb_nmax > 0.05, else 0 (the calibration sigmoid and the
// column-major transpose are fixed transforms and are left to gpb_kernel).
__global__ __launch_bounds__(256) void gpb64_kernel(const double* __restrict__ gcue, const double* __restrict__ spb,
                                                    const double* __restrict__ nmax, size_t N, double wspb,
                                                    double* __restrict__ orient, double* __restrict__ gmax) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < N; p += size_t(gridDim.x) * blockDim.x) {
    double mv = -200.0;
    for (int o = 0; o < kNumOri; o++) {
      const double g = fma(wspb, spb[o * N + p], gcue[o * N + p]);
      if (orient) orient[o * N + p] = g;
      mv = g > mv ? g : mv;
    }
    if (gmax) gmax[p] = nmax[p] > 0.05 ? mv : 0.0;
  }
}

void launch_gpb_f64(Context& c, int W, int H, const double* d_gcue, const double* d_spb, const double* d_nmax,
                    double* d_orient, double* d_gmax) {
  const size_t N = size_t(W) * H;
  KScope ks(c, kKGpb, double(N) * 8 * 25);
  hipLaunchKernelGGL(gpb64_kernel, dim3(std::min<unsigned>(4096, div_up(N, 256))), dim3(256), 0, c.stream, d_gcue,
                     d_spb, d_nmax, N, kGPbWeights[kNumCues], d_orient, d_gmax);
  GPB_LAUNCH_CHECK();
}

}  // namespace gpb

// Texton histogram features of the superpixel classifier (GLSVM), built on the
// stage (a) texton map: per superpixel (SPFeatures::genTextonFeatures,
// src/SPFeatures.cpp:259-272) and per clipped 65 x 65 window around given
// centres (the texton part of SPFeatures::genRecPatchFeatures,
// src/SPFeatures.cpp:478-545).  Each feature is 10 * count / pixels, in FP64,
// divided then scaled as the reference does.
#include "gpb_pipeline.hpp"

namespace gpb {
namespace {

constexpr int kFeatBins = kTextonK;  // 32 textons
constexpr int kRecHalf = 32;         // patchSize / 2 (src/SPFeatures.cpp:491)

// counts[label][texton] += 1 for every pixel (global atomics: the per-label
// sums are small next to the texton stage that precedes them).
__global__ __launch_bounds__(256) void sp_count_kernel(const uint8_t* __restrict__ tex,
                                                       const int32_t* __restrict__ sp, size_t N, int nlabels,
                                                       unsigned* __restrict__ counts, int* __restrict__ bad) {
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < N; p += size_t(gridDim.x) * blockDim.x) {
    const int l = sp[p];
    if (l < 0 || l >= nlabels) {
      atomicOr(bad, 1);
      continue;
    }
    atomicAdd(&counts[size_t(l) * kFeatBins + tex[p]], 1u);
  }
}

// descr[label][t] = counts / size * 10 (0 for a label with no pixels).
__global__ __launch_bounds__(256) void sp_normalise_kernel(const unsigned* __restrict__ counts, int nlabels,
                                                           double* __restrict__ descr) {
  const int l = blockIdx.x * (blockDim.x / kFeatBins) + threadIdx.x / kFeatBins;
  const int t = threadIdx.x % kFeatBins;
  if (l >= nlabels) return;
  unsigned size = 0;
  for (int k = 0; k < kFeatBins; k++) size += counts[size_t(l) * kFeatBins + k];
  double v = double(counts[size_t(l) * kFeatBins + t]);
  if (size > 0) {
    v /= double(size);
    v *= 10.0;
  }
  descr[size_t(l) * kFeatBins + t] = v;
}

// One workgroup per centre (row, col): LDS histogram over the clipped window.
__global__ __launch_bounds__(256) void rec_hist_kernel(const uint8_t* __restrict__ tex, int W, int H,
                                                       const int32_t* __restrict__ centres,
                                                       double* __restrict__ descr) {
  __shared__ unsigned h[kFeatBins];
  const int k = blockIdx.x;
  const int row = centres[2 * k], col = centres[2 * k + 1];
  if (threadIdx.x < kFeatBins) h[threadIdx.x] = 0;
  __syncthreads();
  const int north = min(kRecHalf, row), south = min(kRecHalf, H - 1 - row);
  const int west = min(kRecHalf, col), east = min(kRecHalf, W - 1 - col);
  const int rh = north + south + 1, rw = west + east + 1;
  for (int i = threadIdx.x; i < rh * rw; i += blockDim.x) {
    const int y = row - north + i / rw, x = col - west + i % rw;
    atomicAdd(&h[tex[size_t(y) * W + x]], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kFeatBins) {
    double v = double(h[threadIdx.x]);
    v /= double(rh * rw);
    v *= 10.0;
    descr[size_t(k) * kFeatBins + threadIdx.x] = v;
  }
}

}  // namespace

void launch_texton_features(Context& c, const uint8_t* d_tex, int W, int H, const int32_t* d_sp, int nlabels,
                            double* d_descr) {
  const size_t N = size_t(W) * H;
  unsigned* counts = c.ws<unsigned>("feat.counts", size_t(nlabels) * kFeatBins);
  int* bad = c.ws<int>("feat.bad", 1);
  GPB_HIP(hipMemsetAsync(counts, 0, sizeof(unsigned) * size_t(nlabels) * kFeatBins, c.stream));
  GPB_HIP(hipMemsetAsync(bad, 0, sizeof(int), c.stream));
  hipLaunchKernelGGL(sp_count_kernel, dim3(std::min<unsigned>(4096, div_up(N, 256))), dim3(256), 0, c.stream, d_tex,
                     d_sp, N, nlabels, counts, bad);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(sp_normalise_kernel, dim3(div_up(size_t(nlabels), 256 / kFeatBins)), dim3(256), 0, c.stream,
                     counts, nlabels, d_descr);
  GPB_LAUNCH_CHECK();
  int bad_h = 0;
  GPB_HIP(hipMemcpyAsync(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
  if (bad_h) throw ArgError("superpixel label outside [0, nlabels)");
}

void launch_texton_patch_features(Context& c, const uint8_t* d_tex, int W, int H, const int32_t* d_centres,
                                  int ncentres, double* d_descr) {
  if (ncentres == 0) return;
  hipLaunchKernelGGL(rec_hist_kernel, dim3(ncentres), dim3(256), 0, c.stream, d_tex, W, H, d_centres, d_descr);
  GPB_LAUNCH_CHECK();
}

}  // namespace gpb

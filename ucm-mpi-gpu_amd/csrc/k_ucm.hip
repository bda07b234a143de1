// Per-pixel parts of OWT-UCM on the GPU, for oriented gPb maps that are
// already in HBM (the output of the gPb path): the maximum over orientations,
// its range and its 8-bit scaling (CreateFinestPartition,
// src/contours2ucm.cpp:255-278), and each contour pixel's strength from the
// maps of its edges' orientations (:363-402).  The watershed, contour
// fitting, cleaning and merging between them are sequential host work
// (owt_ucm.cpp); only the 8-bit map (1 B/px) goes down and the orientation
// mask (1 B/px) up, the strength (8 B/px) down.  Every result is bit-identical
// to the host code on the same values: max and min are exact, and the scaling
// keeps the multiply and the add separate (no contraction) before rounding
// half to even.
#include <algorithm>
#include <vector>

#include "gpb_pipeline.hpp"
#include "owt_ucm.hpp"

namespace gpb {
namespace {

constexpr int kUcmBlock = 256;

// pb = max over the 8 maps (FP32 in, FP32 out: exact), and each block's
// min and max of pb.
__global__ __launch_bounds__(kUcmBlock) void ucm_max_kernel(const float* __restrict__ pbo, size_t N,
                                                            float* __restrict__ pb, float* __restrict__ bmin,
                                                            float* __restrict__ bmax) {
  __shared__ float smin[kUcmBlock / 64], smax[kUcmBlock / 64];
  float lo = __builtin_inff(), hi = -__builtin_inff();
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += size_t(gridDim.x) * blockDim.x) {
    float m = pbo[i];
#pragma unroll
    for (int o = 1; o < 8; o++) {
      const float v = pbo[size_t(o) * N + i];
      if (m < v) m = v;
    }
    pb[i] = m;
    lo = fminf(lo, m);
    hi = fmaxf(hi, m);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, off, 64));
    hi = fmaxf(hi, __shfl_xor(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    smin[threadIdx.x >> 6] = lo;
    smax[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kUcmBlock / 64; w++) {
      lo = fminf(lo, smin[w]);
      hi = fmaxf(hi, smax[w]);
    }
    bmin[blockIdx.x] = fminf(smin[0], lo);
    bmax[blockIdx.x] = fmaxf(smax[0], hi);
  }
}

// pb8 = saturate(round_half_even(pb * alpha + beta)) in FP64.
__global__ __launch_bounds__(kUcmBlock) void ucm_scale_kernel(const float* __restrict__ pb, size_t N, double alpha,
                                                              double beta, uint8_t* __restrict__ pb8) {
#pragma clang fp contract(off)
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += size_t(gridDim.x) * blockDim.x) {
    const double t = double(pb[i]) * alpha;
    const double r = rint(t + beta);
    pb8[i] = !(r >= 0.0) ? uint8_t(0) : (r > 255.0 ? uint8_t(255) : uint8_t(r));
  }
}

// ws_wt = max(0, map_o over the orientation bits of the pixel), in FP64.
__global__ __launch_bounds__(kUcmBlock) void ucm_gather_kernel(const float* __restrict__ pbo,
                                                               const uint8_t* __restrict__ omask, size_t N,
                                                               double* __restrict__ ws_wt) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += size_t(gridDim.x) * blockDim.x) {
    const unsigned m = omask[i];
    double v = 0.0;
    if (m) {
#pragma unroll
      for (int o = 0; o < 8; o++)
        if (m & (1u << o)) {
          const double p = double(pbo[size_t(o) * N + i]);
          v = (p < v) ? v : p;
        }
    }
    ws_wt[i] = v;
  }
}

unsigned ucm_grid(size_t N) { return std::max(1u, std::min(div_up(N, kUcmBlock), 8u * 1024u)); }

}  // namespace

void ucm_contours2ucm_device(Context& c, const float* d_pbo, int width, int height, bool double_size, double* ucm,
                             owt::UcmStages* stages) {
  const size_t N = size_t(width) * height;
  const unsigned g = ucm_grid(N);
  float* d_pb = c.ws<float>("ucm.pb", N);
  float* d_bmin = c.ws<float>("ucm.bmin", g);
  float* d_bmax = c.ws<float>("ucm.bmax", g);
  uint8_t* d_pb8 = c.ws<uint8_t>("ucm.pb8", N);
  {
    KScope ks(c, kKUcm, double(N) * (8 * 4 + 4));
    hipLaunchKernelGGL(ucm_max_kernel, dim3(g), dim3(kUcmBlock), 0, c.stream, d_pbo, N, d_pb, d_bmin, d_bmax);
    GPB_LAUNCH_CHECK();
  }
  std::vector<float> bmin(g), bmax(g);
  GPB_HIP(hipMemcpyAsync(bmin.data(), d_bmin, sizeof(float) * g, hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipMemcpyAsync(bmax.data(), d_bmax, sizeof(float) * g, hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
  c.kt.fold();
  const double lo = *std::min_element(bmin.begin(), bmin.end()), hi = *std::max_element(bmax.begin(), bmax.end());
  // the reference's alpha and beta (src/contours2ucm.cpp:272-273)
  const double alpha = 255 / (hi - lo);
  const double beta = -255 * lo / (hi - lo);
  owt::UcmStages local;
  owt::UcmStages& S = stages ? *stages : local;
  S.pb8.resize(N);
  {
    KScope ks(c, kKUcm, double(N) * 5);
    hipLaunchKernelGGL(ucm_scale_kernel, dim3(g), dim3(kUcmBlock), 0, c.stream, d_pb, N, alpha, beta, d_pb8);
    GPB_LAUNCH_CHECK();
  }
  GPB_HIP(hipMemcpyAsync(S.pb8.data(), d_pb8, N, hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
  c.kt.fold();
  std::vector<uint8_t> omask;
  owt::finest_partition(height, width, S, omask);
  uint8_t* d_mask = c.ws<uint8_t>("ucm.mask", N);
  double* d_wt = c.ws<double>("ucm.wswt", N);
  GPB_HIP(hipMemcpyAsync(d_mask, omask.data(), N, hipMemcpyHostToDevice, c.stream));
  {
    // the mask, one map value per contour pixel and orientation (about 2% of
    // the pixels), the strength out
    KScope ks(c, kKUcm, double(N) * (1 + 8));
    hipLaunchKernelGGL(ucm_gather_kernel, dim3(g), dim3(kUcmBlock), 0, c.stream, d_pbo, d_mask, N, d_wt);
    GPB_LAUNCH_CHECK();
  }
  S.ws_wt.resize(N);
  GPB_HIP(hipMemcpyAsync(S.ws_wt.data(), d_wt, sizeof(double) * N, hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
  c.kt.fold();
  owt::ucm_from_strength(height, width, double_size, ucm, S);
}

}  // namespace gpb

// Stage (a): Schmid filter-bank responses and texton assignment, fused.
//
// Reference: SPFeatures::genTextonMap (SPFeatures.cpp:277-360) — grey image,
// REFLECT border of 24, 13 filter2D correlations with the 49x49 Schmid bank,
// then the nearest column of the 13x32 codebook.  Here one workgroup stages a
// 64x64 grey tile (16x16 outputs + 24-pixel halo) in LDS, each lane folds the
// 2401 taps into 273 radial classes (exact integer sums) and accumulates the 13
// responses with one fmaf per class in ascending-radius order, then takes the
// argmin of the FP32 squared distances (fmaf over the 13 features, lowest index
// wins ties).  The responses never touch HBM: 4 bytes in, 1 byte out per pixel.
#include "gpb_device.hpp"

namespace gpb {

constexpr int kTexTX = 16, kTexTY = 16;
constexpr int kTexTileW = kTexTX + 2 * kSchmidHalf;  // 64
constexpr int kTexTileH = kTexTY + 2 * kSchmidHalf;  // 64

__device__ __forceinline__ int d_reflect(int i, int n) {
  if (n == 1) return 0;
  const int p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - 1 - i;
}

__device__ __forceinline__ float d_grey(uint32_t px) {
  const int r = px & 0xff, g = (px >> 8) & 0xff, b = (px >> 16) & 0xff;
  return float((r * 4899 + g * 9617 + b * 1868 + 8192) >> 14);
}

__global__ __launch_bounds__(256) void texton_kernel(const uint32_t* __restrict__ rgb, int W, int H,
                                                     const int* __restrict__ cls_off, const int* __restrict__ cls_start,
                                                     const float* __restrict__ cls_w, int ncls,
                                                     const float* __restrict__ codebook, uint8_t* __restrict__ labels) {
  __shared__ float tile[kTexTileH * kTexTileW];
  const int tid = threadIdx.y * kTexTX + threadIdx.x;
  const int gx0 = blockIdx.x * kTexTX - kSchmidHalf, gy0 = blockIdx.y * kTexTY - kSchmidHalf;
  for (int i = tid; i < kTexTileH * kTexTileW; i += kTexTX * kTexTY) {
    const int ty = i / kTexTileW, tx = i - ty * kTexTileW;
    const int gx = d_reflect(gx0 + tx, W), gy = d_reflect(gy0 + ty, H);
    tile[i] = d_grey(rgb[size_t(gy) * W + gx]);
  }
  __syncthreads();
  const int px = blockIdx.x * kTexTX + threadIdx.x, py = blockIdx.y * kTexTY + threadIdx.y;
  const float* centre = tile + (threadIdx.y + kSchmidHalf) * kTexTileW + threadIdx.x + kSchmidHalf;
  float r[kSchmidNF];
#pragma unroll
  for (int f = 0; f < kSchmidNF; f++) r[f] = 0.f;
  int m = 0;
  for (int c = 0; c < ncls; c++) {
    const int end = cls_start[c + 1];
    // eight members in flight: grey levels are integers and a class sums at
    // most 2401 x 255 < 2^24, so the FP32 partial sums are exact and their
    // order does not change s
    float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (; m + 8 <= end; m += 8) {
#pragma unroll
      for (int u = 0; u < 8; u++) a[u] += centre[cls_off[m + u]];
    }
    for (; m < end; m++) a[0] += centre[cls_off[m]];
    const float s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
#pragma unroll
    for (int f = 0; f < kSchmidNF; f++) r[f] = __builtin_fmaf(cls_w[f * ncls + c], s, r[f]);
  }
  if (px >= W || py >= H) return;
  float best = __builtin_huge_valf();
  int bk = 0;
  for (int k = 0; k < kTextonK; k++) {
    float acc = 0.f;
#pragma unroll
    for (int f = 0; f < kSchmidNF; f++) {
      const float d = r[f] - codebook[f * kTextonK + k];
      acc = __builtin_fmaf(d, d, acc);
    }
    if (acc < best) {
      best = acc;
      bk = k;
    }
  }
  labels[size_t(py) * W + px] = uint8_t(bk);
}

void upload_schmid(Context& c) {
  const SchmidTable& t = schmid_table();
  std::vector<int> off(t.member_dx.size());
  for (size_t m = 0; m < off.size(); m++) off[m] = t.member_dy[m] * kTexTileW + t.member_dx[m];
  c.schmid_off.alloc(off.size());
  c.schmid_start.alloc(t.class_start.size());
  c.schmid_w.alloc(t.weight.size());
  GPB_HIP(hipMemcpy(c.schmid_off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
  GPB_HIP(hipMemcpy(c.schmid_start.p, t.class_start.data(), t.class_start.size() * sizeof(int), hipMemcpyHostToDevice));
  GPB_HIP(hipMemcpy(c.schmid_w.p, t.weight.data(), t.weight.size() * sizeof(float), hipMemcpyHostToDevice));
  c.schmid_classes = t.num_classes;
}

void launch_texton(Context& c, const uint32_t* d_rgb, int W, int H, const float* d_codebook, uint8_t* d_labels) {
  dim3 block(kTexTX, kTexTY);
  dim3 grid(div_up(W, kTexTX), div_up(H, kTexTY));
  KScope ks(c, kKTexton, 5.0 * W * H);
  hipLaunchKernelGGL(texton_kernel, grid, block, 0, c.stream, d_rgb, W, H, c.schmid_off.p, c.schmid_start.p,
                     c.schmid_w.p, c.schmid_classes, d_codebook, d_labels);
  GPB_LAUNCH_CHECK();
}

}  // namespace gpb

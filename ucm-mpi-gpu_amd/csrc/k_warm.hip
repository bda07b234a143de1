// Learned warm start for block Lanczos (no counterpart in the reference; its
// solvers start from a fixed or random vector: spectralPb.cpp:221-244,
// damascene/src/noReorthog/lanczos.cu:393-752).
//
// A 4-level convolutional encoder-decoder over the cue features at 1/4
// resolution predicts 8 vectors per 4x4 pixel cell.  Upsampled to pixels,
// they span an approximation of the 8 nontrivial generalised eigenvectors
// and, with the locked vector D^1/2 1, form the 9-vector starting block of
// block Lanczos.
//
// Network (cells h0 x w0, channels in brackets; every conv is 3x3, pad 1):
//   x0[64] -> a0 = relu(c0)[32] -> s0 = relu(c1)[32]
//   p1 = pool(s0) -> a1 = relu(c2)[64] -> s1 = relu(c3)[64]
//   p2 = pool(s1) -> a2 = relu(c4)[128] -> s2 = relu(c5)[128]
//   p3 = pool(s2) -> a3 = relu(c6)[128] -> b3 = relu(c7)[128]
//   u2 = up(b3) + s2 -> d2 = relu(c8)[64]
//   u1 = up(d2) + s1 -> d1 = relu(c9)[32]
//   u0 = up(d1) + s0 -> d0 = relu(c10)[32] -> y = c11[32] (channels 0..7 used;
//   the head is 32 wide so that its input gradient reduces over 32 channels)
// pool = 2x2 mean, up = 2x nearest.  Activations and gradients are BF16,
// stored NHWC in "padded planes": each image's (h+2) x wp grid (wp = w + 2
// rounded up to 8) with a zero border, so a 3x3 tap is a constant offset of
// the flat position.  Convolutions are implicit GEMMs on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation); the input gradient reuses
// the forward kernel with negated tap offsets and transposed weights; the
// weight gradient is a split-K GEMM over positions on transposed copies.
//
// Loss (per image, cells c with D-weight Db_c = sum of D over the cell and
// targets T_c = sum over the cell of D_p v_p for eigenvectors v_1..v_8):
// P' = P - 1 m^T (D-orthogonal to the constant), G = P'^T Db P',
// C = P'^T T, L = 1 - tr(G^-1 C C^T) / 8, the mean squared sine of the
// principal angles between span(upsampled P) and span(v_1..v_8) in the D
// inner product (invariant to sign and rotation within the subspace).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "k_warm.hpp"

namespace gpb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// BF16 with round-to-nearest-even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// ---- geometry ------------------------------------------------------------------

struct Geo {
  int nb = 0, h = 0, w = 0, hp = 0, wp = 0;
  size_t plane() const { return size_t(hp) * wp; }
  size_t Q() const { return plane() * nb; }
  // readable margin (positions) before and after a tensor of this geometry
  size_t margin() const { return size_t(wp) + 128; }
};
static Geo make_geo(int nb, int h, int w) {
  Geo g;
  g.nb = nb;
  g.h = h;
  g.w = w;
  g.hp = h + 2;
  g.wp = (w + 2 + 7) / 8 * 8;
  return g;
}

struct GeoD {  // device copy
  int hp, wp, h, w;
};
static GeoD dev(const Geo& g) { return GeoD{g.hp, g.wp, g.h, g.w}; }

__device__ __forceinline__ bool interior(long long q, const GeoD& g, int* n, int* y, int* x) {
  const long long plane = (long long)g.hp * g.wp;
  const long long nn = q / plane;
  const long long r = q - nn * plane;
  const int yy = int(r / g.wp), xx = int(r - (long long)yy * g.wp);
  *n = int(nn);
  *y = yy - 1;
  *x = xx - 1;
  return yy >= 1 && yy <= g.h && xx >= 1 && xx <= g.w;
}
__device__ __forceinline__ long long gpos(const GeoD& g, int n, int y, int x) {
  return ((long long)n * g.hp + (y + 1)) * g.wp + (x + 1);
}

// ---- convolution: implicit GEMM on MFMA ---------------------------------------

struct ConvArgs {
  const uint16_t* in;    // position 0; positions -margin .. Q + margin are readable
  const uint16_t* wt;    // [9][cout][cin] BF16
  const float* bias;     // [cout] or null
  void* out;             // position 0, [Q][cout], BF16 or FP32
  const uint16_t* mask;  // optional [Q][cout]: output zeroed where mask <= 0
  int cin, cout;
  long long Q;
  GeoD g;
  int tapsign;  // +1: out[q] = sum_t in[q + d_t] W_t;  -1: in[q - d_t]
  int relu;
};

// Wave: 16 positions x 16*NCO output channels; workgroup: 64 positions.
// MFMA operands: lane l supplies row (l & 15) and the 8 reduction indices
// 8 (l >> 4) .. +7 of A (positions x input channels) and of B (output
// channels x input channels); the accumulator holds rows 4 (l >> 4) + i,
// column l & 15.
template <int NCO, bool F32OUT>
__global__ __launch_bounds__(256) void conv3x3_kernel(ConvArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const long long q0 = (long long)blockIdx.x * 64 + wv * 16;
  const int co0 = blockIdx.y * (NCO * 16);
  f32x4 acc[NCO];
#pragma unroll
  for (int nb = 0; nb < NCO; nb++) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < 9; t++) {
    const int d = a.tapsign * ((t / 3 - 1) * a.g.wp + (t % 3 - 1));
    const uint16_t* arow = a.in + (q0 + row + d) * a.cin + kg * 8;
    const uint16_t* brow = a.wt + (size_t(t) * a.cout + co0 + row) * a.cin + kg * 8;
    for (int cc = 0; cc < a.cin; cc += 32) {
      const bf16x8 av = load8(arow + cc);
      bf16x8 bv[NCO];
#pragma unroll
      for (int nb = 0; nb < NCO; nb++) bv[nb] = load8(brow + size_t(nb) * 16 * a.cin + cc);
#pragma unroll
      for (int nb = 0; nb < NCO; nb++) acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[nb], acc[nb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const long long q = q0 + kg * 4 + i;
    if (q >= a.Q) continue;
    int n, y, x;
    const bool in = interior(q, a.g, &n, &y, &x);
#pragma unroll
    for (int nb = 0; nb < NCO; nb++) {
      const int co = co0 + nb * 16 + row;
      float v = acc[nb][i];
      if (a.bias) v += a.bias[co];
      if (a.relu) v = fmaxf(v, 0.f);
      if (a.mask && !(bf2f(a.mask[q * a.cout + co]) > 0.f)) v = 0.f;
      if (!in) v = 0.f;
      if (F32OUT)
        static_cast<float*>(a.out)[q * a.cout + co] = v;
      else
        static_cast<uint16_t*>(a.out)[q * a.cout + co] = f2bf(v);
    }
  }
}

// ---- weight gradient: split-K GEMM over positions ------------------------------

struct WgradArgs {
  const uint16_t* doT;  // [cout][ld], column MT + q holds dout[q]
  const uint16_t* inT;  // 3 copies (dx = -1, 0, 1), each [cin][ld]: column MT + q holds in[q + dx]
  size_t ld, stride_dx;
  int MT;
  int cin, cout, wp;
  int kchunk;
  long long Kpad;
  float* dW;  // [9][cout][cin], accumulated
};

// Wave: (tap, 16*TM output channels, 16*TN input channels, one chunk of
// kchunk positions); grid.x covers the chunks four per workgroup.
template <int TM, int TN>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const long long k0 = ((long long)blockIdx.x * 4 + wv) * a.kchunk;
  if (k0 >= a.Kpad) return;  // no barriers in this kernel
  const long long k1 = min(k0 + (long long)a.kchunk, a.Kpad);
  const int ntiles = a.cin / (16 * TN), mtiles = a.cout / (16 * TM);
  int yb = blockIdx.y;
  const int nt = yb % ntiles;
  yb /= ntiles;
  const int mt = yb % mtiles;
  const int t = yb / mtiles;
  const int dy = t / 3 - 1, dx = t % 3 - 1;
  const uint16_t* A = a.doT + size_t(mt * 16 * TM + row) * a.ld + a.MT + kg * 8;
  const uint16_t* B = a.inT + size_t(dx + 1) * a.stride_dx + size_t(nt * 16 * TN + row) * a.ld +
                      (long long)a.MT + (long long)dy * a.wp + kg * 8;
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long long k = k0; k < k1; k += 32) {
    bf16x8 av[TM], bv[TN];
#pragma unroll
    for (int i = 0; i < TM; i++) av[i] = load8(A + size_t(i) * 16 * a.ld + k);
#pragma unroll
    for (int j = 0; j < TN; j++) bv[j] = load8(B + size_t(j) * 16 * a.ld + k);
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int co = mt * 16 * TM + i * 16 + kg * 4 + e;
        const int ci = nt * 16 * TN + j * 16 + row;
        atomicAdd(&a.dW[(size_t(t) * a.cout + co) * a.cin + ci], acc[i][j][e]);
      }
}

// T[c][j] = X[j - MT + sh][c] for j < ld (64-position tiles through LDS).
__global__ __launch_bounds__(256) void transpose_kernel(const uint16_t* __restrict__ X, int C,
                                                        uint16_t* __restrict__ T, size_t ld, int MT, int sh) {
  __shared__ uint16_t tile[64][130];
  const long long j0 = (long long)blockIdx.x * 64;
  const int cg8 = C / 8;
  for (int e = threadIdx.x; e < 64 * cg8; e += 256) {
    const int p = e / cg8, cg = e - p * cg8;
    const long long j = j0 + p;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (j < (long long)ld) v = *reinterpret_cast<const uint4*>(X + (j - MT + sh) * C + cg * 8);
    const uint16_t* hv = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
    for (int k = 0; k < 8; k++) tile[p][cg * 8 + k] = hv[k];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < C * 8; e += 256) {
    const int c = e / 8, pg = e - c * 8;
    const long long j = j0 + pg * 8;
    if (j + 8 <= (long long)ld) {
      uint4 v;
      uint16_t* hv = reinterpret_cast<uint16_t*>(&v);
#pragma unroll
      for (int k = 0; k < 8; k++) hv[k] = tile[pg * 8 + k][c];
      *reinterpret_cast<uint4*>(T + size_t(c) * ld + j) = v;
    }
  }
}

// db[c] += sum over columns [MT, MT + K) of row c of T.
__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t* __restrict__ T, size_t ld, int MT, long long K,
                                                        float* __restrict__ db) {
  __shared__ float sh[4];
  const int c = blockIdx.y;
  float s = 0.f;
  for (long long j = (long long)blockIdx.x * 256 + threadIdx.x; j < K; j += (long long)gridDim.x * 256)
    s += bf2f(T[size_t(c) * ld + MT + j]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&db[c], sh[0] + sh[1] + sh[2] + sh[3]);
}

// ---- resampling and gradient routing (8 channels per thread) -------------------

__device__ __forceinline__ void ld8f(const uint16_t* p, float* f) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
  for (int k = 0; k < 8; k++) f[k] = bf2f(h[k]);
}
__device__ __forceinline__ void st8f(uint16_t* p, const float* f) {
  uint4 v;
  uint16_t* h = reinterpret_cast<uint16_t*>(&v);
#pragma unroll
  for (int k = 0; k < 8; k++) h[k] = f2bf(f[k]);
  *reinterpret_cast<uint4*>(p) = v;
}

// out (coarse) = 2x2 mean of in (fine)
__global__ __launch_bounds__(256) void pool_kernel(uint16_t* __restrict__ out, GeoD gc, long long Qc,
                                                   const uint16_t* __restrict__ in, GeoD gf, int C) {
  const int cg8 = C / 8;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < Qc * cg8; e += (long long)gridDim.x * 256) {
    const long long q = e / cg8;
    const int cg = int(e - q * cg8);
    int n, y, x;
    float o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (interior(q, gc, &n, &y, &x)) {
#pragma unroll
      for (int d = 0; d < 4; d++) {
        float f[8];
        ld8f(in + gpos(gf, n, 2 * y + (d >> 1), 2 * x + (d & 1)) * C + cg * 8, f);
#pragma unroll
        for (int k = 0; k < 8; k++) o[k] += f[k];
      }
#pragma unroll
      for (int k = 0; k < 8; k++) o[k] *= 0.25f;
    }
    st8f(out + q * C + cg * 8, o);
  }
}

// out (fine) = up(coarse) + skip (fine)
__global__ __launch_bounds__(256) void upadd_kernel(uint16_t* __restrict__ out, GeoD gf, long long Qf,
                                                    const uint16_t* __restrict__ coarse, GeoD gc,
                                                    const uint16_t* __restrict__ skip, int C) {
  const int cg8 = C / 8;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < Qf * cg8; e += (long long)gridDim.x * 256) {
    const long long q = e / cg8;
    const int cg = int(e - q * cg8);
    int n, y, x;
    float o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (interior(q, gf, &n, &y, &x)) {
      float u[8], s[8];
      ld8f(coarse + gpos(gc, n, y >> 1, x >> 1) * C + cg * 8, u);
      ld8f(skip + q * C + cg * 8, s);
#pragma unroll
      for (int k = 0; k < 8; k++) o[k] = u[k] + s[k];
    }
    st8f(out + q * C + cg * 8, o);
  }
}

// out (coarse) = (2x2 sum of g (fine)) * (mask (coarse) > 0): gradient of a
// nearest upsampling followed by a ReLU mask
__global__ __launch_bounds__(256) void upsum_mask_kernel(uint16_t* __restrict__ out, GeoD gc, long long Qc,
                                                         const uint16_t* __restrict__ g, GeoD gf,
                                                         const uint16_t* __restrict__ mask, int C) {
  const int cg8 = C / 8;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < Qc * cg8; e += (long long)gridDim.x * 256) {
    const long long q = e / cg8;
    const int cg = int(e - q * cg8);
    int n, y, x;
    float o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (interior(q, gc, &n, &y, &x)) {
#pragma unroll
      for (int d = 0; d < 4; d++) {
        float f[8];
        ld8f(g + gpos(gf, n, 2 * y + (d >> 1), 2 * x + (d & 1)) * C + cg * 8, f);
#pragma unroll
        for (int k = 0; k < 8; k++) o[k] += f[k];
      }
      float m[8];
      ld8f(mask + q * C + cg * 8, m);
#pragma unroll
      for (int k = 0; k < 8; k++) o[k] = m[k] > 0.f ? o[k] : 0.f;
    }
    st8f(out + q * C + cg * 8, o);
  }
}

// out (fine) = (gskip (fine) + gpool(coarse) / 4 at the parent cell) * (mask > 0):
// gradient of s where s feeds both a skip connection and a 2x2 mean pool
__global__ __launch_bounds__(256) void combine_mask_kernel(uint16_t* __restrict__ out, GeoD gf, long long Qf,
                                                           const uint16_t* __restrict__ gskip,
                                                           const uint16_t* __restrict__ gpool, GeoD gc,
                                                           const uint16_t* __restrict__ mask, int C) {
  const int cg8 = C / 8;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < Qf * cg8; e += (long long)gridDim.x * 256) {
    const long long q = e / cg8;
    const int cg = int(e - q * cg8);
    int n, y, x;
    float o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (interior(q, gf, &n, &y, &x)) {
      float a[8], b[8], m[8];
      ld8f(gskip + q * C + cg * 8, a);
      ld8f(gpool + gpos(gc, n, y >> 1, x >> 1) * C + cg * 8, b);
      ld8f(mask + q * C + cg * 8, m);
#pragma unroll
      for (int k = 0; k < 8; k++) o[k] = m[k] > 0.f ? a[k] + 0.25f * b[k] : 0.f;
    }
    st8f(out + q * C + cg * 8, o);
  }
}

// ---- inputs, targets, outputs ------------------------------------------------------

constexpr float kWarmFeatScale = 4.f;
constexpr int kHead = 32;  // head output channels (0..7 carry the block)

// x0 of image n from d_feat [32][hq][wq]: channels 0..31 = 4 feat, 32/33 =
// cell coordinates in (-1, 1), the rest 0.
__global__ __launch_bounds__(256) void pack_input_kernel(const float* __restrict__ feat, int hq, int wq,
                                                         uint16_t* __restrict__ x0, GeoD g, int n) {
  const long long total = (long long)g.h * g.w;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int y = int(e / g.w), x = int(e - (long long)y * g.w);
    uint16_t* dst = x0 + gpos(g, n, y, x) * kWarmIn;
    const bool in = y < hq && x < wq;
    for (int c = 0; c < kWarmFeat; c++)
      dst[c] = f2bf(in ? kWarmFeatScale * feat[(size_t(c) * hq + y) * wq + x] : 0.f);
    dst[kWarmFeat] = f2bf((2.f * x + 1.f) / float(wq) - 1.f);
    dst[kWarmFeat + 1] = f2bf((2.f * y + 1.f) / float(hq) - 1.f);
    for (int c = kWarmFeat + 2; c < kWarmIn; c++) dst[c] = 0;
  }
}

// Cell sums of image n: Db[cell] = sum D_p, T[cell][k] = sum D_p v_{k+1}(p)
// over the cell's pixels (eigenvector frame: padded by kICRadius).
__global__ __launch_bounds__(256) void pool_targets_kernel(const float* __restrict__ evecs, size_t Np,
                                                           const float* __restrict__ D, int W, int H, int Wpad,
                                                           int hq, int wq, float* __restrict__ T,
                                                           float* __restrict__ Db, GeoD g, int n) {
  const long long total = (long long)hq * wq;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int cy = int(e / wq), cx = int(e - (long long)cy * wq);
    double sd = 0, st[kWarmOut] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int dy = 0; dy < 4; dy++) {
      const int y = 4 * cy + dy;
      if (y >= H) break;
      for (int dx = 0; dx < 4; dx++) {
        const int x = 4 * cx + dx;
        if (x >= W) break;
        const size_t p = size_t(y + kICRadius) * Wpad + x + kICRadius;
        const double d = D[p];
        sd += d;
#pragma unroll
        for (int k = 0; k < kWarmOut; k++) st[k] += d * double(evecs[size_t(k + 1) * Np + p]);
      }
    }
    const long long q = gpos(g, n, cy, cx);
    Db[q] = float(sd);
#pragma unroll
    for (int k = 0; k < kWarmOut; k++) T[q * kWarmOut + k] = float(st[k]);
  }
}

// Nearest-cell upsampling of the network output y [Q][kHead] (image 0) to the
// eigensolver frame, plus a 1e-3 relative hashed perturbation that keeps the
// block full-rank where cells repeat.
__global__ __launch_bounds__(256) void upsample_block_kernel(const float* __restrict__ y, GeoD g, int W, int H,
                                                             int Wpad, size_t Np, const float* __restrict__ rms,
                                                             float* __restrict__ X) {
  const long long total = (long long)W * H;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int py = int(e / W), px = int(e - (long long)py * W);
    const long long q = gpos(g, 0, py >> 2, px >> 2);
    const size_t p = size_t(py + kICRadius) * Wpad + px + kICRadius;
#pragma unroll
    for (int k = 0; k < kWarmOut; k++) {
      uint64_t z = uint64_t(p) * 0x9E3779B97F4A7C15ull + uint64_t(k + 1) * 0xBF58476D1CE4E5B9ull;
      z ^= z >> 31;
      z *= 0x94D049BB133111EBull;
      z ^= z >> 29;
      const float u = float(z >> 40) * (1.f / 16777216.f) - 0.5f;
      X[size_t(k) * Np + p] = y[q * kHead + k] + 2e-3f * rms[k] * u;
    }
  }
}

// per-channel sum of squares of y channels 0..7 (image 0), for the
// perturbation scale
__global__ __launch_bounds__(256) void channel_sq_kernel(const float* __restrict__ y, long long Q,
                                                         float* __restrict__ acc) {
  float s[kWarmOut] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < Q; q += (long long)gridDim.x * 256)
#pragma unroll
    for (int k = 0; k < kWarmOut; k++) s[k] += y[q * kHead + k] * y[q * kHead + k];
#pragma unroll
  for (int k = 0; k < kWarmOut; k++) {
    float v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&acc[k], v);
  }
}

__global__ void rms_finish_kernel(float* acc, float count) {
  const int k = threadIdx.x;
  if (k < kWarmOut) acc[k] = sqrtf(acc[k] / fmaxf(count, 1.f));
}

// Copies dataset image `src` into batch slot `dst` under one of the 8
// symmetries of the square cell grid (tf bit 0: flip x, bit 1: flip y, bit 2:
// transpose; applied in that order to the output position).  Cue channel
// 8 c + o moves with the line direction: a flip maps the angle o pi/8 to
// -o pi/8, a transpose to pi/2 - o pi/8; the coordinate channels stay with
// the position; targets and weights move with the geometry.
__global__ __launch_bounds__(256) void gather_sym_kernel(const uint16_t* __restrict__ xs, const float* __restrict__ Ts,
                                                         const float* __restrict__ Ds, GeoD g, int src,
                                                         uint16_t* __restrict__ xd, float* __restrict__ Td,
                                                         float* __restrict__ Dd, int dst, int tf) {
  const int total = g.h * g.w;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < total; e += gridDim.x * 256) {
    const int y = e / g.w, x = e - y * g.w;
    int sx = x, sy = y;
    if (tf & 1) sx = g.w - 1 - sx;
    if (tf & 2) sy = g.h - 1 - sy;
    if (tf & 4) {
      const int t = sx;
      sx = sy;
      sy = t;
    }
    const long long qs = gpos(g, src, sy, sx), qd = gpos(g, dst, y, x), qc = gpos(g, src, y, x);
    for (int c = 0; c < kWarmFeat; c++) {
      int o = c & 7;
      if (tf & 1) o = (8 - o) & 7;
      if (tf & 2) o = (8 - o) & 7;
      if (tf & 4) o = (12 - o) & 7;
      xd[qd * kWarmIn + c] = xs[qs * kWarmIn + (c & ~7) + o];
    }
    for (int c = kWarmFeat; c < kWarmIn; c++) xd[qd * kWarmIn + c] = xs[qc * kWarmIn + c];
#pragma unroll
    for (int k = 0; k < kWarmOut; k++) Td[qd * kWarmOut + k] = Ts[qs * kWarmOut + k];
    Dd[qd] = Ds[qs];
  }
}

// ---- loss -----------------------------------------------------------------------

constexpr int kLossStats = 117;  // s, u[8], G[36], C[64], tsum[8]

// partial[(n * nchunk + chunk) * 117 + t]: statistic t over the chunk's
// positions of image n (each thread accumulates one statistic).
__global__ __launch_bounds__(128) void loss_stats_kernel(const float* __restrict__ P, const float* __restrict__ T,
                                                         const float* __restrict__ Db, long long plane, int chunk,
                                                         double* __restrict__ partial) {
  const int t = threadIdx.x;
  if (t >= kLossStats) return;
  const int n = blockIdx.y, nchunk = gridDim.x;
  int kind, ia = 0, ib = 0;
  if (t == 0) {
    kind = 0;
  } else if (t < 9) {
    kind = 1;
    ia = t - 1;
  } else if (t < 45) {
    kind = 2;
    int r = t - 9, a = 0;
    while (r >= 8 - a) {
      r -= 8 - a;
      a++;
    }
    ia = a;
    ib = a + r;
  } else if (t < 109) {
    kind = 3;
    ia = (t - 45) / 8;
    ib = (t - 45) % 8;
  } else {
    kind = 4;
    ia = t - 109;
  }
  const long long lo = (long long)blockIdx.x * chunk, hi = min(lo + chunk, plane);
  double s = 0;
  for (long long j = lo; j < hi; j++) {
    const long long q = (long long)n * plane + j;
    const double d = Db[q];
    switch (kind) {
      case 0: s += d; break;
      case 1: s += d * P[q * kHead + ia]; break;
      case 2: s += d * double(P[q * kHead + ia]) * P[q * kHead + ib]; break;
      case 3: s += double(P[q * kHead + ia]) * T[q * kWarmOut + ib]; break;
      default: s += T[q * kWarmOut + ia]; break;
    }
  }
  partial[(size_t(n) * nchunk + blockIdx.x) * kLossStats + t] = s;
}

// dy[q][k] = (M1 T_q + Db_q (M2 P_q + g0))_k for k < 8, 0 for k < kHead
// (coef per image: M1[64], M2[64], g0[8]).
__global__ __launch_bounds__(256) void loss_grad_kernel(const float* __restrict__ P, const float* __restrict__ T,
                                                        const float* __restrict__ Db, const float* __restrict__ coef,
                                                        long long plane, long long Q, uint16_t* __restrict__ dy) {
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < Q; q += (long long)gridDim.x * 256) {
    const float* cf = coef + (q / plane) * 136;
    float p[8], tt[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      p[k] = P[q * kHead + k];
      tt[k] = T[q * kWarmOut + k];
    }
    const float d = Db[q];
    uint4 v0, v1;
    uint16_t* h0 = reinterpret_cast<uint16_t*>(&v0);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      float a = 0.f, b = cf[128 + k];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        a += cf[k * 8 + j] * tt[j];
        b += cf[64 + k * 8 + j] * p[j];
      }
      h0[k] = f2bf(a + d * b);
    }
    v1 = make_uint4(0, 0, 0, 0);
    *reinterpret_cast<uint4*>(dy + q * kHead) = v0;
#pragma unroll
    for (int k = 8; k < kHead; k += 8) *reinterpret_cast<uint4*>(dy + q * kHead + k) = v1;
  }
}

// ---- optimiser ------------------------------------------------------------------

__global__ __launch_bounds__(256) void sumsq_kernel(const float* __restrict__ g, size_t n, double* __restrict__ acc) {
  double s = 0;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) s += double(g[i]) * g[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(acc, s);
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, size_t n, float lr,
                                                   float b1, float b2, float eps, float bc1, float bc2,
                                                   const double* __restrict__ gnorm2, float clip) {
  const float gn = float(sqrt(*gnorm2));
  const float scale = gn > clip ? clip / gn : 1.f;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const float gi = g[i] * scale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

// BF16 weights of one layer: wf[t][co][ci] (forward) and wd[t][ci][co]
// (input gradient) from the FP32 master W[t][co][ci].
__global__ __launch_bounds__(256) void convert_weights_kernel(const float* __restrict__ W, int cin, int cout,
                                                              uint16_t* __restrict__ wf, uint16_t* __restrict__ wd) {
  const int n = 9 * cin * cout;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int t = i / (cin * cout), r = i - t * cin * cout, co = r / cin, ci = r - co * cin;
    const uint16_t b = f2bf(W[i]);
    wf[i] = b;
    wd[(size_t(t) * cin + ci) * cout + co] = b;
  }
}

__global__ void zero_rows_kernel(float* W, int cin, int cout, int co_from) {
  // zero W[t][co][ci] and bias entries for co >= co_from (bias follows W)
  const int n = 9 * cin * cout;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n + cout; i += gridDim.x * 256) {
    if (i < n) {
      const int co = (i / cin) % cout;
      if (co >= co_from) W[i] = 0.f;
    } else if (i - n >= co_from) {
      W[i] = 0.f;
    }
  }
}

// ---- host side --------------------------------------------------------------------

namespace {

struct LayerSpec {
  int cin, cout, relu;
};
constexpr LayerSpec kLayers[kWarmLayers] = {
    {64, 32, 1},  {32, 32, 1},   {32, 64, 1},   {64, 64, 1},  {64, 128, 1}, {128, 128, 1},
    {128, 128, 1}, {128, 128, 1}, {128, 64, 1}, {64, 32, 1},  {32, 32, 1},  {32, kHead, 0}};

struct LayerOff {
  size_t w, b;
};

const std::vector<LayerOff>& layer_offsets() {
  static const std::vector<LayerOff> off = [] {
    std::vector<LayerOff> o;
    size_t at = 0;
    for (const auto& L : kLayers) {
      LayerOff lo;
      lo.w = at;
      at += size_t(9) * L.cin * L.cout;
      lo.b = at;
      at += L.cout;
      o.push_back(lo);
    }
    return o;
  }();
  return off;
}

size_t param_total() {
  const auto& o = layer_offsets();
  return o.back().b + kLayers[kWarmLayers - 1].cout;
}

struct Ten {
  uint16_t* p = nullptr;  // position 0
  int C = 0;
  Geo g;
};

}  // namespace

struct WarmState {
  bool ready = false;
  int adam_t = 0;
  // training set
  int ds_W = 0, ds_H = 0, ds_n = 0, ds_cap = 0;
  int ds_hq = 0, ds_wq = 0;
  Geo ds_geo;  // nb = capacity
  uint16_t* ds_x = nullptr;
  float* ds_T = nullptr;
  float* ds_Db = nullptr;
};

static WarmState& state(Context& c) {
  if (!c.warm) c.warm = std::make_shared<WarmState>();
  return *c.warm;
}

size_t warm_param_count() { return param_total(); }

void warm_layer_info(int l, int* cin, int* cout, int* relu, size_t* w_off, size_t* b_off) {
  if (l < 0 || l >= kWarmLayers) throw ArgError("layer index out of range");
  const auto& o = layer_offsets()[size_t(l)];
  *cin = kLayers[l].cin;
  *cout = kLayers[l].cout;
  *relu = kLayers[l].relu;
  *w_off = o.w;
  *b_off = o.b;
}

bool warm_ready(const Context& c) { return c.warm && c.warm->ready; }

void warm_cell_grid(int W, int H, int* hq, int* wq, int* h0, int* w0) {
  *hq = (H + 3) / 4;
  *wq = (W + 3) / 4;
  *h0 = (*hq + 7) / 8 * 8;
  *w0 = (*wq + 7) / 8 * 8;
}

namespace {

float* params(Context& c) { return c.ws<float>("warm.param", param_total()); }

void refresh_bf16(Context& c) {
  const size_t n = param_total();
  float* P = params(c);
  uint16_t* wf = c.ws<uint16_t>("warm.wf", n);
  uint16_t* wd = c.ws<uint16_t>("warm.wd", n);
  const auto& off = layer_offsets();
  for (int l = 0; l < kWarmLayers; l++) {
    const int cnt = 9 * kLayers[l].cin * kLayers[l].cout;
    hipLaunchKernelGGL(convert_weights_kernel, dim3(std::min(1024, (cnt + 255) / 256)), dim3(256), 0, c.stream,
                       P + off[size_t(l)].w, kLayers[l].cin, kLayers[l].cout, wf + off[size_t(l)].w,
                       wd + off[size_t(l)].w);
    GPB_LAUNCH_CHECK();
  }
}

// Tensor of geometry g and C channels in workspace `name`, zero-filled
// (margins included).
Ten tensor(Context& c, const std::string& name, const Geo& g, int C) {
  const size_t mg = g.margin();
  const size_t count = (g.Q() + 2 * mg) * size_t(C);
  uint16_t* base = c.ws<uint16_t>(name, count);
  GPB_HIP(hipMemsetAsync(base, 0, count * sizeof(uint16_t), c.stream));
  Ten t;
  t.p = base + mg * C;
  t.C = C;
  t.g = g;
  return t;
}

unsigned grid_for(long long work) { return unsigned(std::min<long long>(4096, std::max<long long>(1, (work + 255) / 256))); }

void conv(Context& c, const Ten& in, int l, bool dgrad, void* out, bool f32out, const Ten* mask) {
  const LayerSpec& L = kLayers[l];
  const size_t off = layer_offsets()[size_t(l)].w;
  ConvArgs a;
  a.in = in.p;
  a.cin = dgrad ? L.cout : L.cin;
  a.cout = dgrad ? L.cin : L.cout;
  if (in.C != a.cin) throw std::logic_error("warm start: channel mismatch");
  a.wt = (dgrad ? c.ws<uint16_t>("warm.wd", param_total()) : c.ws<uint16_t>("warm.wf", param_total())) + off;
  a.bias = dgrad ? nullptr : params(c) + layer_offsets()[size_t(l)].b;
  a.out = out;
  a.mask = mask ? mask->p : nullptr;
  a.Q = (long long)in.g.Q();
  a.g = dev(in.g);
  a.tapsign = dgrad ? -1 : 1;
  a.relu = (!dgrad && L.relu) ? 1 : 0;
  const int nco = std::min(8, a.cout / 16);
  const dim3 grid(div_up(size_t(a.Q), 64), a.cout / (16 * nco));
  KScope ks(c, kKWarm, double(a.Q) * 2.0 * (a.cin + a.cout * (f32out ? 2 : 1)));
  if (f32out) {
    if (nco != 2) throw std::logic_error("warm start: FP32 output only for the head");
    hipLaunchKernelGGL((conv3x3_kernel<2, true>), grid, dim3(256), 0, c.stream, a);
  } else {
    switch (nco) {
      case 1: hipLaunchKernelGGL((conv3x3_kernel<1, false>), grid, dim3(256), 0, c.stream, a); break;
      case 2: hipLaunchKernelGGL((conv3x3_kernel<2, false>), grid, dim3(256), 0, c.stream, a); break;
      case 4: hipLaunchKernelGGL((conv3x3_kernel<4, false>), grid, dim3(256), 0, c.stream, a); break;
      case 8: hipLaunchKernelGGL((conv3x3_kernel<8, false>), grid, dim3(256), 0, c.stream, a); break;
      default: throw std::logic_error("warm start: unsupported channel count");
    }
  }
  GPB_LAUNCH_CHECK();
}

// dW, db of layer l from its input and the gradient at its output (after the
// ReLU mask), both at the layer's resolution.
void wgrad(Context& c, const Ten& in, const Ten& dout, int l, float* grad) {
  const LayerSpec& L = kLayers[l];
  const Geo& g = in.g;
  const int MT = g.wp;
  const long long Kpad = (long long)((g.Q() + 31) / 32 * 32);
  const size_t ld = (size_t(2 * MT) + size_t(Kpad) + 63) / 64 * 64;
  uint16_t* doT = c.ws<uint16_t>("warm.doT", size_t(L.cout) * ld);
  uint16_t* inT = c.ws<uint16_t>("warm.inT", size_t(3) * L.cin * ld);
  const unsigned tg = div_up(ld, 64);
  hipLaunchKernelGGL(transpose_kernel, dim3(tg), dim3(256), 0, c.stream, dout.p, L.cout, doT, ld, MT, 0);
  GPB_LAUNCH_CHECK();
  for (int dx = -1; dx <= 1; dx++) {
    hipLaunchKernelGGL(transpose_kernel, dim3(tg), dim3(256), 0, c.stream, in.p, L.cin,
                       inT + size_t(dx + 1) * L.cin * ld, ld, MT, dx);
    GPB_LAUNCH_CHECK();
  }
  const auto& off = layer_offsets()[size_t(l)];
  WgradArgs a;
  a.doT = doT;
  a.inT = inT;
  a.ld = ld;
  a.stride_dx = size_t(L.cin) * ld;
  a.MT = MT;
  a.cin = L.cin;
  a.cout = L.cout;
  a.wp = g.wp;
  a.kchunk = 1024;
  a.Kpad = Kpad;
  a.dW = grad + off.w;
  const int tm = std::min(4, L.cout / 16), tn = std::min(4, L.cin / 16);
  const long long chunks = (Kpad + a.kchunk - 1) / a.kchunk;
  const dim3 grid(unsigned((chunks + 3) / 4), unsigned(9 * (L.cout / (16 * tm)) * (L.cin / (16 * tn))));
  KScope ks(c, kKWarm, double(Kpad) * 2.0 * (L.cin * 4.0 + L.cout * 2.0));
#define GPB_WGRAD(TM, TN) hipLaunchKernelGGL((wgrad_kernel<TM, TN>), grid, dim3(256), 0, c.stream, a)
  if (tm == 1 && tn == 2) GPB_WGRAD(1, 2);
  else if (tm == 2 && tn == 2) GPB_WGRAD(2, 2);
  else if (tm == 2 && tn == 4) GPB_WGRAD(2, 4);
  else if (tm == 4 && tn == 2) GPB_WGRAD(4, 2);
  else if (tm == 4 && tn == 4) GPB_WGRAD(4, 4);
  else throw std::logic_error("warm start: unsupported weight-gradient tile");
#undef GPB_WGRAD
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(bias_grad_kernel, dim3(std::min<unsigned>(256, div_up(size_t(Kpad), 256)), L.cout), dim3(256), 0,
                     c.stream, doT, ld, MT, Kpad, grad + off.b);
  GPB_LAUNCH_CHECK();
}

struct Net {
  Geo g0, g1, g2, g3;
  Ten x0, a0, s0, p1, a1, s1, p2, a2, s2, p3, a3, b3, u2, d2, u1, d1, u0, d0;
  float* y = nullptr;  // [Q0][kHead] FP32 (position 0)
};

void launch_pool(Context& c, const Ten& out, const Ten& in) {
  const long long work = (long long)out.g.Q() * (out.C / 8);
  hipLaunchKernelGGL(pool_kernel, dim3(grid_for(work)), dim3(256), 0, c.stream, out.p, dev(out.g),
                     (long long)out.g.Q(), in.p, dev(in.g), out.C);
  GPB_LAUNCH_CHECK();
}
void launch_upadd(Context& c, const Ten& out, const Ten& coarse, const Ten& skip) {
  const long long work = (long long)out.g.Q() * (out.C / 8);
  hipLaunchKernelGGL(upadd_kernel, dim3(grid_for(work)), dim3(256), 0, c.stream, out.p, dev(out.g),
                     (long long)out.g.Q(), coarse.p, dev(coarse.g), skip.p, out.C);
  GPB_LAUNCH_CHECK();
}
void launch_upsum_mask(Context& c, const Ten& out, const Ten& g, const Ten& mask) {
  const long long work = (long long)out.g.Q() * (out.C / 8);
  hipLaunchKernelGGL(upsum_mask_kernel, dim3(grid_for(work)), dim3(256), 0, c.stream, out.p, dev(out.g),
                     (long long)out.g.Q(), g.p, dev(g.g), mask.p, out.C);
  GPB_LAUNCH_CHECK();
}
void launch_combine_mask(Context& c, const Ten& out, const Ten& gskip, const Ten& gpool, const Ten& mask) {
  const long long work = (long long)out.g.Q() * (out.C / 8);
  hipLaunchKernelGGL(combine_mask_kernel, dim3(grid_for(work)), dim3(256), 0, c.stream, out.p, dev(out.g),
                     (long long)out.g.Q(), gskip.p, gpool.p, dev(gpool.g), mask.p, out.C);
  GPB_LAUNCH_CHECK();
}

// Forward pass; the input x0 (64 channels at g0) must be filled by the caller
// after this call returns the tensor (net.x0 is allocated here unless given).
void alloc_forward(Context& c, Net& n, int nb, int h0, int w0, const uint16_t* x0_external) {
  n.g0 = make_geo(nb, h0, w0);
  n.g1 = make_geo(nb, h0 / 2, w0 / 2);
  n.g2 = make_geo(nb, h0 / 4, w0 / 4);
  n.g3 = make_geo(nb, h0 / 8, w0 / 8);
  if (x0_external) {
    n.x0.p = const_cast<uint16_t*>(x0_external);
    n.x0.C = kWarmIn;
    n.x0.g = n.g0;
  } else {
    n.x0 = tensor(c, "warm.x0", n.g0, kWarmIn);
  }
  n.a0 = tensor(c, "warm.a0", n.g0, 32);
  n.s0 = tensor(c, "warm.s0", n.g0, 32);
  n.p1 = tensor(c, "warm.p1", n.g1, 32);
  n.a1 = tensor(c, "warm.a1", n.g1, 64);
  n.s1 = tensor(c, "warm.s1", n.g1, 64);
  n.p2 = tensor(c, "warm.p2", n.g2, 64);
  n.a2 = tensor(c, "warm.a2", n.g2, 128);
  n.s2 = tensor(c, "warm.s2", n.g2, 128);
  n.p3 = tensor(c, "warm.p3", n.g3, 128);
  n.a3 = tensor(c, "warm.a3", n.g3, 128);
  n.b3 = tensor(c, "warm.b3", n.g3, 128);
  n.u2 = tensor(c, "warm.u2", n.g2, 128);
  n.d2 = tensor(c, "warm.d2", n.g2, 64);
  n.u1 = tensor(c, "warm.u1", n.g1, 64);
  n.d1 = tensor(c, "warm.d1", n.g1, 32);
  n.u0 = tensor(c, "warm.u0", n.g0, 32);
  n.d0 = tensor(c, "warm.d0", n.g0, 32);
  n.y = c.ws<float>("warm.y", n.g0.Q() * kHead);
}

void forward(Context& c, Net& n) {
  conv(c, n.x0, 0, false, n.a0.p, false, nullptr);
  conv(c, n.a0, 1, false, n.s0.p, false, nullptr);
  launch_pool(c, n.p1, n.s0);
  conv(c, n.p1, 2, false, n.a1.p, false, nullptr);
  conv(c, n.a1, 3, false, n.s1.p, false, nullptr);
  launch_pool(c, n.p2, n.s1);
  conv(c, n.p2, 4, false, n.a2.p, false, nullptr);
  conv(c, n.a2, 5, false, n.s2.p, false, nullptr);
  launch_pool(c, n.p3, n.s2);
  conv(c, n.p3, 6, false, n.a3.p, false, nullptr);
  conv(c, n.a3, 7, false, n.b3.p, false, nullptr);
  launch_upadd(c, n.u2, n.b3, n.s2);
  conv(c, n.u2, 8, false, n.d2.p, false, nullptr);
  launch_upadd(c, n.u1, n.d2, n.s1);
  conv(c, n.u1, 9, false, n.d1.p, false, nullptr);
  launch_upadd(c, n.u0, n.d1, n.s0);
  conv(c, n.u0, 10, false, n.d0.p, false, nullptr);
  conv(c, n.d0, 11, false, n.y, true, nullptr);
}

// Backward pass from dy (kHead channels at g0, BF16); accumulates into grad.
void backward(Context& c, Net& n, const Ten& dy, float* grad) {
  Ten g_d0 = tensor(c, "warm.g_d0", n.g0, 32);
  Ten g_u0 = tensor(c, "warm.g_u0", n.g0, 32);
  Ten g_d1 = tensor(c, "warm.g_d1", n.g1, 32);
  Ten g_u1 = tensor(c, "warm.g_u1", n.g1, 64);
  Ten g_d2 = tensor(c, "warm.g_d2", n.g2, 64);
  Ten g_u2 = tensor(c, "warm.g_u2", n.g2, 128);
  Ten g_b3 = tensor(c, "warm.g_b3", n.g3, 128);
  Ten g_a3 = tensor(c, "warm.g_a3", n.g3, 128);
  Ten g_p3 = tensor(c, "warm.g_p3", n.g3, 128);
  Ten g_s2 = tensor(c, "warm.g_s2", n.g2, 128);
  Ten g_a2 = tensor(c, "warm.g_a2", n.g2, 128);
  Ten g_p2 = tensor(c, "warm.g_p2", n.g2, 64);
  Ten g_s1 = tensor(c, "warm.g_s1", n.g1, 64);
  Ten g_a1 = tensor(c, "warm.g_a1", n.g1, 64);
  Ten g_p1 = tensor(c, "warm.g_p1", n.g1, 32);
  Ten g_s0 = tensor(c, "warm.g_s0", n.g0, 32);
  Ten g_a0 = tensor(c, "warm.g_a0", n.g0, 32);
  wgrad(c, n.d0, dy, 11, grad);
  conv(c, dy, 11, true, g_d0.p, false, &n.d0);
  wgrad(c, n.u0, g_d0, 10, grad);
  conv(c, g_d0, 10, true, g_u0.p, false, nullptr);
  launch_upsum_mask(c, g_d1, g_u0, n.d1);
  wgrad(c, n.u1, g_d1, 9, grad);
  conv(c, g_d1, 9, true, g_u1.p, false, nullptr);
  launch_upsum_mask(c, g_d2, g_u1, n.d2);
  wgrad(c, n.u2, g_d2, 8, grad);
  conv(c, g_d2, 8, true, g_u2.p, false, nullptr);
  launch_upsum_mask(c, g_b3, g_u2, n.b3);
  wgrad(c, n.a3, g_b3, 7, grad);
  conv(c, g_b3, 7, true, g_a3.p, false, &n.a3);
  wgrad(c, n.p3, g_a3, 6, grad);
  conv(c, g_a3, 6, true, g_p3.p, false, nullptr);
  launch_combine_mask(c, g_s2, g_u2, g_p3, n.s2);
  wgrad(c, n.a2, g_s2, 5, grad);
  conv(c, g_s2, 5, true, g_a2.p, false, &n.a2);
  wgrad(c, n.p2, g_a2, 4, grad);
  conv(c, g_a2, 4, true, g_p2.p, false, nullptr);
  launch_combine_mask(c, g_s1, g_u1, g_p2, n.s1);
  wgrad(c, n.a1, g_s1, 3, grad);
  conv(c, g_s1, 3, true, g_a1.p, false, &n.a1);
  wgrad(c, n.p1, g_a1, 2, grad);
  conv(c, g_a1, 2, true, g_p1.p, false, nullptr);
  launch_combine_mask(c, g_s0, g_u0, g_p1, n.s0);
  wgrad(c, n.a0, g_s0, 1, grad);
  conv(c, g_s0, 1, true, g_a0.p, false, &n.a0);
  wgrad(c, n.x0, g_a0, 0, grad);
}

// Per-image loss from the statistics; fills coef (M1, M2, g0 scaled by
// 1/nb) and returns the mean loss over the images.
double loss_from_stats(const std::vector<double>& st, int nb, std::vector<float>& coef) {
  coef.assign(size_t(nb) * 136, 0.f);
  double total = 0;
  for (int n = 0; n < nb; n++) {
    const double* s = st.data() + size_t(n) * kLossStats;
    const double S = std::max(s[0], 1e-300);
    double u[8], G[64], C[64], ts[8];
    for (int k = 0; k < 8; k++) u[k] = s[1 + k];
    int t = 9;
    for (int a = 0; a < 8; a++)
      for (int b = a; b < 8; b++) {
        G[a * 8 + b] = G[b * 8 + a] = s[t] - u[a] * u[b] / S;
        t++;
      }
    for (int k = 0; k < 8; k++) ts[k] = s[109 + k];
    for (int a = 0; a < 8; a++)  // C' = P'^T T with P' = P - 1 u^T / S
      for (int b = 0; b < 8; b++) C[a * 8 + b] = s[45 + a * 8 + b] - u[a] * ts[b] / S;
    // ridge for a (near-)singular Gram matrix
    double tr = 0;
    for (int k = 0; k < 8; k++) tr += G[k * 8 + k];
    for (int k = 0; k < 8; k++) G[k * 8 + k] += 1e-9 * tr / 8 + 1e-300;
    // Gi = G^-1 by Gauss-Jordan
    double Aug[8][16];
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 16; j++) Aug[i][j] = j < 8 ? G[i * 8 + j] : (j - 8 == i ? 1.0 : 0.0);
    for (int col = 0; col < 8; col++) {
      int piv = col;
      for (int r = col + 1; r < 8; r++)
        if (std::fabs(Aug[r][col]) > std::fabs(Aug[piv][col])) piv = r;
      for (int j = 0; j < 16; j++) std::swap(Aug[col][j], Aug[piv][j]);
      const double d = Aug[col][col];
      for (int j = 0; j < 16; j++) Aug[col][j] /= d;
      for (int r = 0; r < 8; r++)
        if (r != col) {
          const double f = Aug[r][col];
          for (int j = 0; j < 16; j++) Aug[r][j] -= f * Aug[col][j];
        }
    }
    double Gi[64], A[64], GiC[64], B[64], tmp[64];
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) Gi[i * 8 + j] = Aug[i][8 + j];
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) {
        double a = 0, g = 0;
        for (int k = 0; k < 8; k++) {
          a += C[i * 8 + k] * C[j * 8 + k];
          g += Gi[i * 8 + k] * C[k * 8 + j];
        }
        A[i * 8 + j] = a;
        GiC[i * 8 + j] = g;
      }
    double f = 0;
    for (int i = 0; i < 8; i++)
      for (int k = 0; k < 8; k++) f += Gi[i * 8 + k] * A[k * 8 + i];
    total += 1.0 - f / 8.0;
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) {
        double a = 0;
        for (int k = 0; k < 8; k++) a += Gi[i * 8 + k] * A[k * 8 + j];
        tmp[i * 8 + j] = a;
      }
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) {
        double a = 0;
        for (int k = 0; k < 8; k++) a += tmp[i * 8 + k] * Gi[k * 8 + j];
        B[i * 8 + j] = a;
      }
    // dL/dP_c = M1 T_c + Db_c (M2 P_c + g0) with M1 = -(2/8) Gi C,
    // M2 = (2/8) B, g0 = -M2 m - M1 tsum / S, m = u / S; scaled by 1/nb
    const double sc = 1.0 / nb;
    float* cf = coef.data() + size_t(n) * 136;
    double M1[64], M2[64];
    for (int i = 0; i < 64; i++) {
      M1[i] = -0.25 * GiC[i];
      M2[i] = 0.25 * B[i];
    }
    for (int k = 0; k < 8; k++) {
      double g = 0;
      for (int j = 0; j < 8; j++) g += -M2[k * 8 + j] * (u[j] / S) - M1[k * 8 + j] * ts[j] / S;
      cf[128 + k] = float(g * sc);
    }
    for (int i = 0; i < 64; i++) {
      cf[i] = float(M1[i] * sc);
      cf[64 + i] = float(M2[i] * sc);
    }
  }
  return total / nb;
}

// Loss statistics of the network output n.y against (T, Db) laid out like
// the output; returns the mean loss; with dy non-null also its gradient.
double loss_and_grad(Context& c, const Net& n, const float* T, const float* Db, Ten* dy) {
  const int nb = n.g0.nb;
  const long long plane = (long long)n.g0.plane();
  const int chunk = 2048;
  const int nchunk = int((plane + chunk - 1) / chunk);
  double* part = c.ws<double>("warm.lstat", size_t(nb) * nchunk * kLossStats);
  hipLaunchKernelGGL(loss_stats_kernel, dim3(nchunk, nb), dim3(128), 0, c.stream, n.y, T, Db, plane, chunk, part);
  GPB_LAUNCH_CHECK();
  std::vector<double> hp(size_t(nb) * nchunk * kLossStats), st(size_t(nb) * kLossStats, 0.0);
  GPB_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
  for (int i = 0; i < nb; i++)
    for (int ch = 0; ch < nchunk; ch++)
      for (int t = 0; t < kLossStats; t++)
        st[size_t(i) * kLossStats + t] += hp[(size_t(i) * nchunk + ch) * kLossStats + t];
  std::vector<float> coef;
  const double loss = loss_from_stats(st, nb, coef);
  if (dy) {
    float* dcoef = c.ws<float>("warm.coef", coef.size());
    GPB_HIP(hipMemcpyAsync(dcoef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(loss_grad_kernel, dim3(grid_for((long long)n.g0.Q())), dim3(256), 0, c.stream, n.y, T, Db,
                       dcoef, plane, (long long)n.g0.Q(), dy->p);
    GPB_LAUNCH_CHECK();
  }
  return loss;
}

void adam_step(Context& c, float* grad, float lr) {
  WarmState& w = state(c);
  const size_t np = param_total();
  double* gn = c.ws<double>("warm.gnorm", 1);
  GPB_HIP(hipMemsetAsync(gn, 0, sizeof(double), c.stream));
  hipLaunchKernelGGL(sumsq_kernel, dim3(256), dim3(256), 0, c.stream, grad, np, gn);
  GPB_LAUNCH_CHECK();
  w.adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.f - std::pow(b1, float(w.adam_t)), bc2 = 1.f - std::pow(b2, float(w.adam_t));
  hipLaunchKernelGGL(adam_kernel, dim3(std::min<unsigned>(1024, div_up(np, 256))), dim3(256), 0, c.stream, params(c),
                     grad, c.ws<float>("warm.m", np), c.ws<float>("warm.v", np), np, lr, b1, b2, 1e-8f, bc1, bc2, gn,
                     1.0f);
  GPB_LAUNCH_CHECK();
  // the head's unused output channels stay zero
  const auto& off = layer_offsets()[kWarmLayers - 1];
  hipLaunchKernelGGL(zero_rows_kernel, dim3(64), dim3(256), 0, c.stream, params(c) + off.w,
                     kLayers[kWarmLayers - 1].cin, kLayers[kWarmLayers - 1].cout, kWarmOut);
  GPB_LAUNCH_CHECK();
  refresh_bf16(c);
}

}  // namespace

void warm_init(Context& c, uint64_t seed) {
  WarmState& w = state(c);
  const size_t np = param_total();
  std::vector<float> h(np, 0.f);
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto next = [&z]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
  };
  const auto& off = layer_offsets();
  for (int l = 0; l < kWarmLayers; l++) {
    const LayerSpec& L = kLayers[l];
    // He initialisation (Box-Muller normals), a smaller scale for the head
    const double sd = (l == kWarmLayers - 1) ? std::sqrt(1.0 / (9.0 * L.cin)) : std::sqrt(2.0 / (9.0 * L.cin));
    for (size_t i = 0; i < size_t(9) * L.cin * L.cout; i++) {
      const double u1 = (double(next() >> 11) + 1.0) / 9007199254740993.0;
      const double u2 = double(next() >> 11) / 9007199254740992.0;
      h[off[size_t(l)].w + i] = float(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
  }
  warm_set_params(c, h.data());
}

void warm_set_params(Context& c, const float* host) {
  WarmState& w = state(c);
  const size_t np = param_total();
  std::vector<float> h(host, host + np);
  // the head's output channels 8..15 carry no target: kept at zero
  const auto& o = layer_offsets()[kWarmLayers - 1];
  const LayerSpec& L = kLayers[kWarmLayers - 1];
  for (int t = 0; t < 9; t++)
    for (int co = kWarmOut; co < L.cout; co++)
      for (int ci = 0; ci < L.cin; ci++) h[o.w + (size_t(t) * L.cout + co) * L.cin + ci] = 0.f;
  for (int co = kWarmOut; co < L.cout; co++) h[o.b + co] = 0.f;
  GPB_HIP(hipMemcpyAsync(params(c), h.data(), np * sizeof(float), hipMemcpyHostToDevice, c.stream));
  GPB_HIP(hipMemsetAsync(c.ws<float>("warm.m", np), 0, np * sizeof(float), c.stream));
  GPB_HIP(hipMemsetAsync(c.ws<float>("warm.v", np), 0, np * sizeof(float), c.stream));
  refresh_bf16(c);
  GPB_HIP(hipStreamSynchronize(c.stream));
  w.adam_t = 0;
  w.ready = true;
}

void warm_get_params(Context& c, float* host) {
  if (!warm_ready(c)) throw ArgError("warm-start network not initialised");
  GPB_HIP(hipMemcpyAsync(host, params(c), param_total() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
}

void warm_predict_block(Context& c, const float* d_feat, int W, int H, float* d_block, size_t Np) {
  if (!warm_ready(c)) throw ArgError("warm-start network not initialised");
  int hq, wq, h0, w0;
  warm_cell_grid(W, H, &hq, &wq, &h0, &w0);
  Net n;
  alloc_forward(c, n, 1, h0, w0, nullptr);
  hipLaunchKernelGGL(pack_input_kernel, dim3(grid_for((long long)h0 * w0)), dim3(256), 0, c.stream, d_feat, hq, wq,
                     n.x0.p, dev(n.g0), 0);
  GPB_LAUNCH_CHECK();
  forward(c, n);
  float* rms = c.ws<float>("warm.rms", kWarmOut);
  GPB_HIP(hipMemsetAsync(rms, 0, kWarmOut * sizeof(float), c.stream));
  hipLaunchKernelGGL(channel_sq_kernel, dim3(grid_for((long long)n.g0.Q())), dim3(256), 0, c.stream, n.y,
                     (long long)n.g0.Q(), rms);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(rms_finish_kernel, dim3(1), dim3(64), 0, c.stream, rms, float(hq) * wq);
  GPB_LAUNCH_CHECK();
  launch_fill(c, d_block, size_t(kWarmOut) * Np, 0.f);
  hipLaunchKernelGGL(upsample_block_kernel, dim3(grid_for((long long)W * H)), dim3(256), 0, c.stream, n.y, dev(n.g0), W,
                     H, W + 2 * kICRadius, Np, rms, d_block);
  GPB_LAUNCH_CHECK();
}

void warm_dataset_reset(Context& c, int W, int H, int capacity) {
  if (W <= 0 || H <= 0 || capacity <= 0) throw ArgError("dataset size must be positive");
  WarmState& w = state(c);
  int hq, wq, h0, w0;
  warm_cell_grid(W, H, &hq, &wq, &h0, &w0);
  w.ds_W = W;
  w.ds_H = H;
  w.ds_hq = hq;
  w.ds_wq = wq;
  w.ds_n = 0;
  w.ds_cap = capacity;
  w.ds_geo = make_geo(capacity, h0, w0);
  const size_t mg = w.ds_geo.margin();
  const size_t Q = w.ds_geo.Q();
  uint16_t* xb = c.ws<uint16_t>("warm.ds_x", (Q + 2 * mg) * kWarmIn);
  GPB_HIP(hipMemsetAsync(xb, 0, (Q + 2 * mg) * kWarmIn * sizeof(uint16_t), c.stream));
  w.ds_x = xb + mg * kWarmIn;
  w.ds_T = c.ws<float>("warm.ds_T", Q * kWarmOut);
  w.ds_Db = c.ws<float>("warm.ds_Db", Q);
  GPB_HIP(hipMemsetAsync(w.ds_T, 0, Q * kWarmOut * sizeof(float), c.stream));
  GPB_HIP(hipMemsetAsync(w.ds_Db, 0, Q * sizeof(float), c.stream));
}

void warm_dataset_add(Context& c, const float* d_feat, const Stencil& st, const float* d_evecs) {
  WarmState& w = state(c);
  if (w.ds_cap == 0) throw ArgError("training set not initialised");
  if (w.ds_n >= w.ds_cap) throw ArgError("training set full");
  if (st.W != w.ds_W || st.H != w.ds_H) throw ArgError("training images must all have the set's size");
  const GeoD g = dev(w.ds_geo);
  const int n = w.ds_n;
  hipLaunchKernelGGL(pack_input_kernel, dim3(grid_for((long long)g.h * g.w)), dim3(256), 0, c.stream, d_feat, w.ds_hq,
                     w.ds_wq, w.ds_x, g, n);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(pool_targets_kernel, dim3(grid_for((long long)w.ds_hq * w.ds_wq)), dim3(256), 0, c.stream, d_evecs,
                     st.Np, st.D, st.W, st.H, st.Wp, w.ds_hq, w.ds_wq, w.ds_T, w.ds_Db, g, n);
  GPB_LAUNCH_CHECK();
  w.ds_n++;
}

int warm_dataset_size(const Context& c) { return c.warm ? c.warm->ds_n : 0; }

namespace {

// Copies dataset images idx[0..nb) into the batch tensors (planes are
// contiguous, so a batch is nb consecutive planes); with tf non-null, image i
// goes through grid symmetry tf[i] (gather_sym_kernel).
void gather_batch(Context& c, const WarmState& w, const int* idx, int nb, const Net& n, float* T, float* Db,
                  const int* tf = nullptr) {
  const size_t plane = w.ds_geo.plane();
  if (tf) {
    const GeoD g = dev(w.ds_geo);
    for (int i = 0; i < nb; i++) {
      hipLaunchKernelGGL(gather_sym_kernel, dim3(grid_for((long long)g.h * g.w)), dim3(256), 0, c.stream, w.ds_x,
                         w.ds_T, w.ds_Db, g, idx[i], n.x0.p, T, Db, i, tf[i]);
      GPB_LAUNCH_CHECK();
    }
    return;
  }
  for (int i = 0; i < nb; i++) {
    const size_t src = size_t(idx[i]) * plane;
    GPB_HIP(hipMemcpyAsync(n.x0.p + size_t(i) * plane * kWarmIn, w.ds_x + src * kWarmIn,
                           plane * kWarmIn * sizeof(uint16_t), hipMemcpyDeviceToDevice, c.stream));
    GPB_HIP(hipMemcpyAsync(T + size_t(i) * plane * kWarmOut, w.ds_T + src * kWarmOut, plane * kWarmOut * sizeof(float),
                           hipMemcpyDeviceToDevice, c.stream));
    GPB_HIP(hipMemcpyAsync(Db + size_t(i) * plane, w.ds_Db + src, plane * sizeof(float), hipMemcpyDeviceToDevice,
                           c.stream));
  }
}

}  // namespace

void warm_train(Context& c, int first, int count, int epochs, int batch, float lr, uint64_t seed, int augment,
                float* epoch_loss) {
  WarmState& w = state(c);
  if (!w.ready) throw ArgError("warm-start network not initialised");
  if (first < 0 || count < 1 || first + count > w.ds_n) throw ArgError("dataset range out of bounds");
  if (epochs < 1 || batch < 1) throw ArgError("epochs and batch must be positive");
  if (augment && w.ds_geo.h != w.ds_geo.w) throw ArgError("symmetry augmentation needs square training images");
  batch = std::min(batch, count);
  Net n;
  alloc_forward(c, n, batch, w.ds_geo.h, w.ds_geo.w, nullptr);
  const size_t plane = n.g0.plane();
  float* T = c.ws<float>("warm.bT", size_t(batch) * plane * kWarmOut);
  float* Db = c.ws<float>("warm.bDb", size_t(batch) * plane);
  // the symmetry gather writes the image positions only: the padding ring
  // the loss also reads must hold zero targets and weights
  GPB_HIP(hipMemsetAsync(T, 0, size_t(batch) * plane * kWarmOut * sizeof(float), c.stream));
  GPB_HIP(hipMemsetAsync(Db, 0, size_t(batch) * plane * sizeof(float), c.stream));
  Ten dy = tensor(c, "warm.dy", n.g0, kHead);
  float* grad = c.ws<float>("warm.grad", param_total());
  std::vector<int> order(static_cast<size_t>(count));
  std::iota(order.begin(), order.end(), first);
  uint64_t z = seed ^ 0xD1B54A32D192ED03ull;
  auto rnd = [&z]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
  };
  for (int e = 0; e < epochs; e++) {
    for (size_t i = order.size(); i > 1; i--) std::swap(order[i - 1], order[size_t(rnd() % i)]);
    double sum = 0;
    int steps = 0;
    std::vector<int> tf(static_cast<size_t>(batch));
    for (int b0 = 0; b0 + batch <= count; b0 += batch) {
      for (auto& t : tf) t = augment ? int(rnd() & 7) : 0;
      gather_batch(c, w, order.data() + b0, batch, n, T, Db, augment ? tf.data() : nullptr);
      forward(c, n);
      sum += loss_and_grad(c, n, T, Db, &dy);
      GPB_HIP(hipMemsetAsync(grad, 0, param_total() * sizeof(float), c.stream));
      backward(c, n, dy, grad);
      adam_step(c, grad, lr);
      steps++;
    }
    GPB_HIP(hipStreamSynchronize(c.stream));
    c.kt.fold();
    const double mean = steps ? sum / steps : 0.0;
    if (epoch_loss) epoch_loss[e] = float(mean);
    if (c.verbose) std::fprintf(stderr, "gpb: warm-start training epoch %d: loss %.5f\n", e + 1, mean);
  }
}

double warm_eval_loss(Context& c, int first, int count) {
  WarmState& w = state(c);
  if (!w.ready) throw ArgError("warm-start network not initialised");
  if (first < 0 || count < 1 || first + count > w.ds_n) throw ArgError("dataset range out of bounds");
  const int batch = std::min(count, 16);
  Net n;
  alloc_forward(c, n, batch, w.ds_geo.h, w.ds_geo.w, nullptr);
  const size_t plane = n.g0.plane();
  float* T = c.ws<float>("warm.bT", size_t(batch) * plane * kWarmOut);
  float* Db = c.ws<float>("warm.bDb", size_t(batch) * plane);
  double sum = 0;
  int done = 0;
  std::vector<int> idx(static_cast<size_t>(batch));
  while (done < count) {
    const int nb = std::min(batch, count - done);
    for (int i = 0; i < batch; i++) idx[size_t(i)] = first + done + std::min(i, nb - 1);
    gather_batch(c, w, idx.data(), batch, n, T, Db);
    forward(c, n);
    // per-image losses: recompute the mean over the first nb images only
    const long long plane_l = (long long)plane;
    const int chunk = 2048;
    const int nchunk = int((plane_l + chunk - 1) / chunk);
    double* part = c.ws<double>("warm.lstat", size_t(batch) * nchunk * kLossStats);
    hipLaunchKernelGGL(loss_stats_kernel, dim3(nchunk, batch), dim3(128), 0, c.stream, n.y, T, Db, plane_l, chunk,
                       part);
    GPB_LAUNCH_CHECK();
    std::vector<double> hp(size_t(batch) * nchunk * kLossStats), st(size_t(nb) * kLossStats, 0.0);
    GPB_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    for (int i = 0; i < nb; i++)
      for (int ch = 0; ch < nchunk; ch++)
        for (int t = 0; t < kLossStats; t++)
          st[size_t(i) * kLossStats + t] += hp[(size_t(i) * nchunk + ch) * kLossStats + t];
    std::vector<float> coef;
    sum += loss_from_stats(st, nb, coef) * nb;
    done += nb;
  }
  c.kt.fold();
  return sum / count;
}

double warm_loss_grad_host(Context& c, int nimg, int hq, int wq, const float* x, const float* T, const float* Db,
                           float* grad, float* y) {
  if (!warm_ready(c)) throw ArgError("warm-start network not initialised");
  if (nimg < 1 || hq < 8 || wq < 8 || hq % 8 || wq % 8) throw ArgError("cell grid must be a positive multiple of 8");
  Net n;
  alloc_forward(c, n, nimg, hq, wq, nullptr);
  const Geo& g = n.g0;
  const size_t Q = g.Q();
  std::vector<uint16_t> hx(Q * kWarmIn, 0);
  std::vector<float> hT(Q * kWarmOut, 0.f), hD(Q, 0.f);
  for (int i = 0; i < nimg; i++)
    for (int yy = 0; yy < hq; yy++)
      for (int xx = 0; xx < wq; xx++) {
        const size_t q = (size_t(i) * g.hp + yy + 1) * g.wp + xx + 1;
        const size_t s = (size_t(i) * hq + yy) * wq + xx;
        for (int ch = 0; ch < kWarmFeat + 2; ch++) hx[q * kWarmIn + ch] = f2bf(x[s * (kWarmFeat + 2) + ch]);
        for (int k = 0; k < kWarmOut; k++) hT[q * kWarmOut + k] = T[s * kWarmOut + k];
        hD[q] = Db[s];
      }
  GPB_HIP(hipMemcpyAsync(n.x0.p, hx.data(), hx.size() * sizeof(uint16_t), hipMemcpyHostToDevice, c.stream));
  float* dT = c.ws<float>("warm.hT", Q * kWarmOut);
  float* dD = c.ws<float>("warm.hD", Q);
  GPB_HIP(hipMemcpyAsync(dT, hT.data(), hT.size() * sizeof(float), hipMemcpyHostToDevice, c.stream));
  GPB_HIP(hipMemcpyAsync(dD, hD.data(), hD.size() * sizeof(float), hipMemcpyHostToDevice, c.stream));
  forward(c, n);
  Ten dy = tensor(c, "warm.dy", g, kHead);
  const double loss = loss_and_grad(c, n, dT, dD, &dy);
  float* dgrad = c.ws<float>("warm.grad", param_total());
  GPB_HIP(hipMemsetAsync(dgrad, 0, param_total() * sizeof(float), c.stream));
  backward(c, n, dy, dgrad);
  if (grad) GPB_HIP(hipMemcpyAsync(grad, dgrad, param_total() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  std::vector<float> hy;
  if (y) {
    hy.resize(Q * kHead);
    GPB_HIP(hipMemcpyAsync(hy.data(), n.y, hy.size() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  }
  GPB_HIP(hipStreamSynchronize(c.stream));
  c.kt.fold();
  if (y)
    for (int i = 0; i < nimg; i++)
      for (int yy = 0; yy < hq; yy++)
        for (int xx = 0; xx < wq; xx++) {
          const size_t q = (size_t(i) * g.hp + yy + 1) * g.wp + xx + 1;
          const size_t s = (size_t(i) * hq + yy) * wq + xx;
          for (int k = 0; k < kWarmOut; k++) y[s * kWarmOut + k] = hy[q * kHead + k];
        }
  return loss;
}

}  // namespace gpb

// Device code shared by the eigensolvers (single-vector, block and batched
// Lanczos): constants, the stencil SpMV tile and block reductions.
#pragma once

#include <vector>

#include "gpb_pipeline.hpp"

namespace gpb {

void symeig_desc(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& z);

constexpr int kRedThreads = 256;
constexpr int kMaxBasis = 128;

constexpr int kProjectMaxN = 64;  // widest basis the fused projection handles

struct SpmvOffsets {
  int o[kStencilHalf];
};

// Normalised Laplacian in difference form: with u = s .* x (s = D^-1/2),
// y_p = s_p * sum_q W_pq (u_p - u_q).  Differences of neighbouring values are
// formed before the products, so eigenvalues near 0 keep their relative
// accuracy in FP32 (the form I - D^-1/2 W D^-1/2 would lose it to cancellation
// against 1).  W is read once per unordered pair: W[k][p] for q = p + e_k and
// W[k][p - e_k] for q = p - e_k.  A 64 x NY*ROWS tile of u (plus a 5-pixel
// halo) is staged in LDS; W streams from HBM with coalesced row reads.
// Each of the NY wavefronts takes ROWS consecutive rows of the tile, and the
// weights are loaded plane group by plane group (stencil_rows): the shifted
// reads W[k][p - e_k] of rows e_k.y above then fall on lines the same wave
// requests for its own rows at about the same time.  On the W access pattern
// alone at 2048^2 (a load-only probe) this took 135 us against 169 us for one
// row per wave with all 80 loads of a pixel together; 129 us reads the 40
// direct planes with no shifted reads at all.
// The shape is picked per image by timing (tune_spmv).
constexpr int kSpmvTX = 64;
constexpr int kSpmvTileW = kSpmvTX + 2 * kICRadius;

struct SpmvArgs {
  int o[kStencilHalf];   // padded linear offsets
  int lo[kStencilHalf];  // LDS tile offsets
};



// One 64 x NY*ROWS tile of y = L x at interior origin (x0, y0) (blockDim =
// 64 x NY); returns this thread's share of x . y over the tile.  AT is the arithmetic type of the
// differences, products and sum (float, or double with the FP32 operands
// widened exactly); W, x, s and y stay FP32 in HBM either way.  u: LDS tile
// of (NY*ROWS+10) x 74 values of s .* x in AT.
// Pointers known to address device global memory (HBM).  Loads through
// them are global_load instructions; through a generic pointer read from a
// table they would be flat loads, which also count against the LDS wait
// counter and serialise against the tile's LDS reads.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<const T> as_global(const T* p) {
  return (gptr<const T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<T> as_global_mut(T* p) {
  return (gptr<T>)p;
}

// Plane-group size of stencil_rows: the loads of G planes for all ROWS rows
// (2 G ROWS floats of registers) are issued together, then their products.
// The group loop stays rolled: unrolled, the compiler hoisted later groups'
// loads and the LDS reads of u, spilled, and ran 1.1-3x slower.
template <int ROWS>
constexpr int spmv_group() {
  return ROWS >= 8 ? 5 : (ROWS >= 4 ? 10 : 20);
}

template <int ROWS, int G>
__device__ __forceinline__ void stencil_load_group(const float* __restrict__ wup_, size_t Np, int Wp, int g,
                                                   const uint32_t (&p4)[ROWS], float (&wa)[G][ROWS],
                                                   float (&wb)[G][ROWS]) {
#pragma unroll
  for (int jj = 0; jj < G; jj++) {
    const int j = g + jj;
    const gptr<const char> plane = as_global(reinterpret_cast<const char*>(wup_ + size_t(j) * Np));
    const uint32_t o4 = uint32_t(kUpper.dy[j] * Wp + kUpper.dx[j]) * 4u;
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      wa[jj][r] = *reinterpret_cast<gptr<const float>>(plane + p4[r]);
      wb[jj][r] = *reinterpret_cast<gptr<const float>>(plane + (p4[r] - o4));
    }
  }
}

template <int ROWS, int G, typename AT>
__device__ __forceinline__ void stencil_fma_group(int g, const int (&c)[ROWS], const AT (&up)[ROWS],
                                                  const AT* __restrict__ u, const float (&wa)[G][ROWS],
                                                  const float (&wb)[G][ROWS], AT (&sum)[ROWS]) {
#pragma unroll
  for (int jj = 0; jj < G; jj++) {
    const int lo = kUpper.dy[g + jj] * kSpmvTileW + kUpper.dx[g + jj];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      sum[r] = fma_t(AT(wa[jj][r]), up[r] - u[c[r] + lo], sum[r]);
      sum[r] = fma_t(AT(wb[jj][r]), up[r] - u[c[r] - lo], sum[r]);
    }
  }
}

// sum[r] = sum_k W[k][p_r] (up_r - u[c_r + lo_k]) + W[k][p_r - e_k] (up_r - u[c_r - lo_k])
// for ROWS pixels (p4: byte offsets within a W plane; c: LDS tile indices),
// terms added in stencil order k = 0..39, the forward term before the
// backward one, whatever the grouping.
template <int ROWS, typename AT>
__device__ __forceinline__ void stencil_rows(const float* __restrict__ wup_, size_t Np, int Wp,
                                             const uint32_t (&p4)[ROWS], const int (&c)[ROWS], const AT (&up)[ROWS],
                                             const AT* __restrict__ u, AT (&sum)[ROWS]) {
  constexpr int G = spmv_group<ROWS>();
  static_assert(kStencilHalf % G == 0, "plane groups must tile the stencil");
#pragma unroll
  for (int r = 0; r < ROWS; r++) sum[r] = 0;
#pragma unroll 1
  for (int g = 0; g < kStencilHalf; g += G) {
    float wa[G][ROWS], wb[G][ROWS];
    stencil_load_group<ROWS, G>(wup_, Np, Wp, g, p4, wa, wb);
    stencil_fma_group<ROWS, G, AT>(g, c, up, u, wa, wb, sum);
  }
}

template <int ROWS, typename AT, int NY = 4>
__device__ __forceinline__ double spmv_tile(const float* __restrict__ wup_, const float* __restrict__ s_,
                                            const float* __restrict__ x_, float* __restrict__ y_, int W, int H, int Wp,
                                            int Hp, size_t Np, const SpmvArgs& /* offsets: kUpper and Wp */, int x0,
                                            int y0, AT* __restrict__ u) {
  const gptr<const float> s = as_global(s_);
  const gptr<const float> x = as_global(x_);
  const gptr<float> y = as_global_mut(y_);
  constexpr int kSpmvTY = NY * ROWS, kSpmvTileH = kSpmvTY + 2 * kICRadius;
  const int tid = threadIdx.y * kSpmvTX + threadIdx.x;
  for (int i = tid; i < kSpmvTileH * kSpmvTileW; i += kSpmvTX * NY) {
    const int ty = i / kSpmvTileW, tx = i - ty * kSpmvTileW;
    const int py = y0 + ty, px = x0 + tx;  // padded coordinates (tile starts at interior - 5)
    AT v = 0;
    if (py < Hp && px < Wp) {
      const size_t q = size_t(py) * Wp + px;
      v = AT(s[q]) * AT(x[q]);
    }
    u[i] = v;
  }
  __syncthreads();
  const int xi = x0 + threadIdx.x;
  double xy = 0;
  if (xi < W) {
    // rows threadIdx.y * ROWS + r of the tile; rows past the image read the
    // weights of its last row (in bounds) and are not written
    uint32_t p4[ROWS];
    int c[ROWS];
    AT up[ROWS], acc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const int ty = threadIdx.y * ROWS + r;
      const int yi = min(y0 + ty, H - 1);
      // 32-bit pixel index: each plane is addressed as a wave-uniform base
      // plus this offset, so the loads need no per-lane 64-bit address
      // arithmetic (planes hold < 2^30 pixels)
      p4[r] = (uint32_t(yi + kICRadius) * uint32_t(Wp) + uint32_t(xi + kICRadius)) * 4u;
      c[r] = (ty + kICRadius) * kSpmvTileW + threadIdx.x + kICRadius;
      up[r] = u[c[r]];
    }
    stencil_rows<ROWS, AT>(wup_, Np, Wp, p4, c, up, u, acc);
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      if (y0 + int(threadIdx.y) * ROWS + r >= H) break;
      const uint32_t p = p4[r] >> 2;
      const float yp = float(AT(s[p]) * acc[r]);
      y[p] = yp;
      xy += double(x[p]) * double(yp);
    }
  }
  return xy;
}

// One Lanczos step of the partially reorthogonalised solver in one pass over
// a tile, the SpMV with the vector updates around it fused in:
//   XMODE 1: x = fl32((w_in - a v_prev) * ib) formed while the tile is staged
//            (the previous step's three-term update and scaling), written to
//            the basis slot v_out for the tile's interior;
//   XMODE 0: x = v_out as stored (a step resumed from the basis);
//   w_out = fl32(L x - b v_prev) (this step's three-term update up to the
//   alpha term, which the next step applies), and the thread's shares of
//   x . w_out, w_out . w_out and x . x in acc[0..2].
// a, ib: the previous step's alpha (0 after a whole-basis pass) and 1/beta;
// b: that beta.  w_in and w_out are distinct buffers (neighbouring tiles read
// w_in in their halos while this tile writes w_out).  Zero halos stay zero:
// only interior pixels are written.  xs: LDS tile of the x values (float).
// The rows of one fused step for a staged window (u, xs: (NY*ROWS + 10) x 74
// values whose row 5 is interior row y0): w_out = fl32(L x - b v_prev) for
// rows y0 .. y0 + NY*ROWS - 1 and the thread's shares of x . w, w . w, x . x
// added to acc.
template <int ROWS, typename AT, int NY>
__device__ __forceinline__ void lanczos_step_rows(const float* __restrict__ wup_, const float* __restrict__ s_,
                                                  const float* __restrict__ v_prev_, float* __restrict__ w_out_,
                                                  double b, int W, int H, int Wp, size_t Np, int x0, int y0,
                                                  const AT* __restrict__ u, const float* __restrict__ xs,
                                                  double acc[3]) {
  const gptr<const float> s = as_global(s_);
  const gptr<const float> v_prev = as_global(v_prev_);
  const gptr<float> w_out = as_global_mut(w_out_);
  const int xi = x0 + threadIdx.x;
  if (xi >= W) return;
  uint32_t p4[ROWS];
  int c[ROWS];
  AT up[ROWS], sum[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; r++) {
    const int ty = threadIdx.y * ROWS + r;
    const int yi = min(y0 + ty, H - 1);  // rows past the image: in-bounds reads, not written
    p4[r] = (uint32_t(yi + kICRadius) * uint32_t(Wp) + uint32_t(xi + kICRadius)) * 4u;
    c[r] = (ty + kICRadius) * kSpmvTileW + threadIdx.x + kICRadius;
    up[r] = u[c[r]];
  }
  stencil_rows<ROWS, AT>(wup_, Np, Wp, p4, c, up, u, sum);
#pragma unroll
  for (int r = 0; r < ROWS; r++) {
    if (y0 + int(threadIdx.y) * ROWS + r >= H) break;
    const uint32_t p = p4[r] >> 2;
    const float vp = v_prev[p];
    const float yp = float(AT(s[p]) * sum[r]);
    const float wv = float(double(yp) - b * double(vp));
    w_out[p] = wv;
    const double xd = double(xs[c[r]]), wd = double(wv);
    acc[0] += xd * wd;
    acc[1] += wd * wd;
    acc[2] += xd * xd;
  }
}

// x and u = s x at one padded pixel of a step's window: XMODE 1 forms
// x = fl32((w_in - a v_prev) * ib), XMODE 0 takes the stored basis vector;
// outside the padded image both are 0.
template <typename AT, int XMODE>
__device__ __forceinline__ void lanczos_stage_value(float wv, float vv, float sv, double a, double ib, float& xv,
                                                    AT& uv) {
  xv = XMODE == 1 ? float((double(wv) - a * double(vv)) * ib) : wv;
  uv = AT(sv) * AT(xv);
}

// One Lanczos step of the partially reorthogonalised solver in one pass over
// a tile, the SpMV with the vector updates around it fused in:
//   XMODE 1: x = fl32((w_in - a v_prev) * ib) formed while the tile is staged
//            (the previous step's three-term update and scaling), written to
//            the basis slot v_out for the tile's interior;
//   XMODE 0: x = v_out as stored (a step resumed from the basis);
//   w_out = fl32(L x - b v_prev) (this step's three-term update up to the
//   alpha term, which the next step applies), and the thread's shares of
//   x . w_out, w_out . w_out and x . x in acc[0..2].
// a, ib: the previous step's alpha (0 after a whole-basis pass) and 1/beta;
// b: that beta.  w_in and w_out are distinct buffers (neighbouring tiles read
// w_in in their halos while this tile writes w_out).  Zero halos stay zero:
// only interior pixels are written.  xs: LDS tile of the x values (float).
template <int ROWS, typename AT, int NY, int XMODE>
__device__ __forceinline__ void lanczos_step_tile(const float* __restrict__ wup_, const float* __restrict__ s_,
                                                  const float* __restrict__ w_in_, const float* __restrict__ v_prev_,
                                                  float* __restrict__ v_out_, float* __restrict__ w_out_, double a,
                                                  double ib, double b, int W, int H, int Wp, int Hp, size_t Np,
                                                  int x0, int y0, AT* __restrict__ u, float* __restrict__ xs,
                                                  double acc[3]) {
  const gptr<const float> s = as_global(s_);
  const gptr<const float> w_in = as_global(w_in_);
  const gptr<const float> v_prev = as_global(v_prev_);
  const gptr<float> v_out = as_global_mut(v_out_);
  constexpr int kSpmvTY = NY * ROWS, kSpmvTileH = kSpmvTY + 2 * kICRadius;
  const int tid = threadIdx.y * kSpmvTX + threadIdx.x;
  for (int i = tid; i < kSpmvTileH * kSpmvTileW; i += kSpmvTX * NY) {
    const int ty = i / kSpmvTileW, tx = i - ty * kSpmvTileW;
    const int py = y0 + ty, px = x0 + tx;  // padded coordinates (tile starts at interior - 5)
    float xv = 0.f;
    AT uv = 0;
    if (py < Hp && px < Wp) {
      const size_t q = size_t(py) * Wp + px;
      lanczos_stage_value<AT, XMODE>(XMODE == 1 ? w_in[q] : v_out[q], XMODE == 1 ? v_prev[q] : 0.f, s[q], a, ib, xv,
                                     uv);
      // the tile's own interior pixels go to the basis
      if (XMODE == 1 && ty >= kICRadius && ty < kICRadius + kSpmvTY && tx >= kICRadius &&
          tx < kICRadius + kSpmvTX && py - kICRadius < H && px - kICRadius < W)
        v_out[q] = xv;
    }
    u[i] = uv;
    xs[i] = xv;
  }
  __syncthreads();
  acc[0] = acc[1] = acc[2] = 0.0;
  lanczos_step_rows<ROWS, AT, NY>(wup_, s_, v_prev_, w_out_, b, W, H, Wp, Np, x0, y0, u, xs, acc);
}

// Block sum of one double per thread over NW wavefronts (blockDim.x * y =
// 64 * NW); every thread gets the result.
template <int NW>
__device__ __forceinline__ double block_sum_waves(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) t += red[w];
  return t;
}

// init + a[0] b[0] + ... + a[n-1] b[n-1], accumulated in order of the index
// (the same rounding as the plain loop), with the loads of 32 terms issued
// before their FMAs: the plain loop waits one memory latency per term, which
// made the one-block recurrence kernels latency-bound.
__device__ __forceinline__ double dot_in_order(const double* __restrict__ a, const double* __restrict__ b, int n,
                                               double init) {
  constexpr int U = 32;
  double s = init;
  for (int l0 = 0; l0 < n; l0 += U) {
    double x[U], y[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      x[u] = l0 + u < n ? a[l0 + u] : 0.0;
      y[u] = l0 + u < n ? b[l0 + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (l0 + u < n) s += x[u] * y[u];
  }
  return s;
}

// Per-thread strided sums s[k] = sum of P[k * G + b] over b = tid, tid + NT, ...
// (in that order) for K arrays of G partials, with eight strides' loads issued
// together.
template <int K, int NT>
__device__ __forceinline__ void strided_sums(const double* __restrict__ P, int G, int tid, double* s) {
  constexpr int U = 8;
#pragma unroll
  for (int k = 0; k < K; k++) s[k] = 0.0;
  for (int b0 = tid; b0 < G; b0 += U * NT) {
    double x[U][K];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int k = 0; k < K; k++) x[u][k] = b0 + u * NT < G ? P[size_t(k) * G + b0 + u * NT] : 0.0;
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int k = 0; k < K; k++)
        if (b0 + u * NT < G) s[k] += x[u][k];
  }
}

// Block sum of one double per thread (256 threads); result in thread 0.
__device__ __forceinline__ double block_sum256(double v, double* red4) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  if ((tid & 63) == 0) red4[tid >> 6] = v;
  __syncthreads();
  return red4[0] + red4[1] + red4[2] + red4[3];
}

}  // namespace gpb

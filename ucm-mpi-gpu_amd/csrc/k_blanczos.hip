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
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gpb_pipeline.hpp"

namespace gpb {

void symeig_desc(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& z);

constexpr int kBMax = 16;
constexpr int kBTX = 32, kBTY = 8;
constexpr int kBTileW = kBTX + 2 * kICRadius, kBTileH = kBTY + 2 * kICRadius;
constexpr int kBMaxBasis = 336;
// Second local Gram-Schmidt pass of a semi-orthogonal block step when a
// column keeps less than eta = 0.01 of its norm through the first: one pass
// leaves local overlaps of about eps_32 ||P|| / ||P'|| <= 100 eps_32 = 6e-6,
// far below the semi-orthogonality tolerance sqrt(eps_32) = 2.4e-4 that the
// next step's measuring pass enforces.
constexpr double kBDgksEta2 = 1e-4;
// CholQR's second pass runs when the column-scaled R of the first has a
// Frobenius condition number of at least this (eps_64 * 1e6 stays far below
// sqrt(eps_32), with room for the FP64 summation error of the Gram).
constexpr double kBChol2Cond = 1e3;

struct BSpmvArgs {
  int o[kStencilHalf];
  int lo[kStencilHalf];
};

// Y_r = L X_r for r < B; X, Y columns are ldx / ldy apart (padded layout).
// u = s x is staged in LDS in FP32 (LDS bandwidth bounds this kernel); AT is
// the accumulation type: with AT = double each product of an FP32 weight and
// an FP32 difference is exact and the 80-term sums carry no FP32 rounding.
// Tiles are dealt XCD-aware in band-major order (as the single-vector SpMV):
// workgroup b runs on XCD b % 8, each XCD takes a contiguous run of tiles
// numbered by vertical bands `band` tiles wide, row-major within a band.
template <int B, typename AT>
__global__ __launch_bounds__(kBTX * kBTY) void bspmv_kernel(const float* __restrict__ wup, const float* __restrict__ s,
                                                            const float* __restrict__ X, size_t ldx,
                                                            float* __restrict__ Y, size_t ldy, int W, int H, int Wp,
                                                            int Hp, size_t Np, BSpmvArgs A, int band) {
  __shared__ float u[kBTileH * kBTileW * B];
  const int tid = threadIdx.y * kBTX + threadIdx.x;
  const int tiles_x = (W + kBTX - 1) / kBTX, tiles_y = (H + kBTY - 1) / kBTY, ntiles = tiles_x * tiles_y;
  const int per_xcd = (ntiles + kNumXcd - 1) / kNumXcd;
  const int tile = int(blockIdx.x % kNumXcd) * per_xcd + int(blockIdx.x / kNumXcd);
  if (tile >= ntiles) return;  // whole workgroup
  const int bw = min(band, tiles_x);
  const int bi = tile / (bw * tiles_y), r0 = tile - bi * bw * tiles_y;
  const int w_here = min(bw, tiles_x - bi * bw);
  const int x0 = (bi * bw + r0 % w_here) * kBTX, y0 = (r0 / w_here) * kBTY;
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
  const uint32_t p = uint32_t(yi + kICRadius) * uint32_t(Wp) + uint32_t(xi + kICRadius);
  const uint32_t p4 = p * 4u;  // byte offset within a plane (wave-uniform plane bases, as in spmv_tile)
  const int c = (threadIdx.y + kICRadius) * kBTileW + threadIdx.x + kICRadius;
  float up[B];
  AT acc[B];
#pragma unroll
  for (int r = 0; r < B; r++) {
    up[r] = u[c * B + r];
    acc[r] = AT(0);
  }
#pragma unroll
  for (int j = 0; j < kStencilHalf; j++) {
    const char* plane = reinterpret_cast<const char*>(wup + size_t(j) * Np);
    const uint32_t o4 = uint32_t(kUpper.dy[j] * Wp + kUpper.dx[j]) * 4u;
    const AT w1 = *reinterpret_cast<const float*>(plane + p4), w2 = *reinterpret_cast<const float*>(plane + (p4 - o4));
    const int lo = kUpper.dy[j] * kBTileW + kUpper.dx[j];
    const float* u1 = u + (c + lo) * B;
    const float* u2 = u + (c - lo) * B;
#pragma unroll
    for (int r = 0; r < B; r++) {
      acc[r] = fma_t(w1, AT(up[r] - u1[r]), acc[r]);
      acc[r] = fma_t(w2, AT(up[r] - u2[r]), acc[r]);
    }
  }
  const AT sp = s[p];
#pragma unroll
  for (int r = 0; r < B; r++) Y[size_t(r) * ldy + p] = float(sp * acc[r]);
}

// Wave-level transpose-reduce of K (a power of two, <= 64) values: every
// step exchanges half of the remaining values with the partner lane, so the
// 64 lanes finish with lane l holding the wave total of value l >> (6 - log2 K)
// after K - 1 + (6 - log2 K) shuffles instead of 6 K.
template <int K, int S>
struct TransposeReduce {
  static __device__ __forceinline__ void run(double* a, int lane) {
    constexpr int h = K / 2;
    const bool up = (lane & S) != 0;
#pragma unroll
    for (int j = 0; j < h; j++) {
      const double send = up ? a[j] : a[j + h];
      const double keep = up ? a[j + h] : a[j];
      a[j] = keep + __shfl_xor(send, S, 64);
    }
    TransposeReduce<h, S / 2>::run(a, lane);
  }
};
template <int S>
struct TransposeReduce<1, S> {
  static __device__ __forceinline__ void run(double* a, int lane) {
#pragma unroll
    for (int s = S; s >= 1; s >>= 1) a[0] += __shfl_xor(a[0], s, 64);
  }
};

constexpr int pow2_ceil(int v) { return v <= 1 ? 1 : 2 * pow2_ceil((v + 1) / 2); }
constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// Panel dots H[i][r] = sum_p V_i[p] (D[p]) W_r[p] for i < n, r < B.  Each
// thread keeps the panel values of PX pixels in registers and streams the
// basis past them in groups of NV vectors (each basis vector is read once per
// launch, the panel once); group totals are transpose-reduced per wave and
// accumulated per workgroup in LDS.  Workgroup g covers chunks g, g + G, ...
// of 256 * PX pixels and writes partial[(i * B + r) * G + g].
//
// dots_chunk is one chunk's share given the panel values w (and weights dd,
// 0 past Np); the fused kernels below call it on values they have just
// computed, with the same chunk mapping, so their sums are the ones a
// separate bdots_kernel launch would produce.
template <int B, int NV, int PX>
struct DotsAcc {
  static constexpr int K = pow2_ceil(NV * B);
  static_assert(K <= 64, "group too wide");
  static constexpr int SH = 6 - ilog2(K);
};
// SELF: the rows are the panel itself (n = B), taken from w instead of
// memory; the loop over row groups is then unrolled.
template <int B, int NV, int PX, bool SELF>
constexpr int kDotsUnroll = SELF ? (B + NV - 1) / NV : 1;
template <int B, int NV, int PX, bool SELF = false>
__device__ __forceinline__ void dots_chunk(const float* V, size_t ld, int n, const size_t (&pp)[PX],
                                           const float (&w)[PX][B], const double (&dd)[PX],
                                           double (*red)[DotsAcc<B, NV, PX>::K], double* tot, int tid) {
  constexpr int K = DotsAcc<B, NV, PX>::K, SH = DotsAcc<B, NV, PX>::SH;
  const int lane = tid & 63, wv = tid >> 6;
  if (SELF) n = B;
#pragma unroll kDotsUnroll<B, NV, PX, SELF>
  for (int i0 = 0; i0 < n; i0 += NV) {
    double a[K];
#pragma unroll
    for (int t = 0; t < K; t++) a[t] = 0.0;
    float x[PX][NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const size_t iv = size_t(min(i0 + v, n - 1));
#pragma unroll
      for (int k = 0; k < PX; k++) {
        if constexpr (SELF) x[k][v] = w[k][min(i0 + v, B - 1)];
        else x[k][v] = V[iv * ld + pp[k]];
      }
    }
#pragma unroll
    for (int k = 0; k < PX; k++)
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const double xv = double(x[k][v]) * dd[k];
#pragma unroll
        for (int r = 0; r < B; r++) a[v * B + r] = fma(xv, double(w[k][r]), a[v * B + r]);
      }
    TransposeReduce<K, 32>::run(a, lane);
    if ((lane & ((1 << SH) - 1)) == 0) red[wv][lane >> SH] = a[0];
    __syncthreads();
    if (tid < NV * B && i0 + tid / B < n) tot[i0 * B + tid] += red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
    __syncthreads();
  }
}

template <int B, int NV, int PX>
__global__ __launch_bounds__(256) void bdots_kernel(const float* __restrict__ V, size_t ld, int n,
                                                    const float* __restrict__ Wm, size_t ldw,
                                                    const float* __restrict__ D, size_t Np,
                                                    double* __restrict__ partial, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;  // whole grid (the gated reduce skips too)
  __shared__ double red[4][DotsAcc<B, NV, PX>::K];
  __shared__ double tot[kBMaxBasis * B];
  const int tid = threadIdx.x;
  const int G = gridDim.x;
  for (int t = tid; t < n * B; t += 256) tot[t] = 0.0;
  __syncthreads();
  const size_t nchunk = (Np + 256 * PX - 1) / (256 * PX);
  for (size_t ch = blockIdx.x; ch < nchunk; ch += G) {
    size_t pp[PX];
    float w[PX][B];
    double dd[PX];
#pragma unroll
    for (int k = 0; k < PX; k++) {
      const size_t p = ch * (256 * PX) + size_t(k) * 256 + tid;
      const bool in = p < Np;
      pp[k] = in ? p : Np - 1;
      dd[k] = in ? (D ? double(D[pp[k]]) : 1.0) : 0.0;
#pragma unroll
      for (int r = 0; r < B; r++) w[k][r] = Wm[size_t(r) * ldw + pp[k]];
    }
    dots_chunk<B, NV, PX>(V, ld, n, pp, w, dd, red, tot, tid);
  }
  for (int t = tid; t < n * B; t += 256) partial[size_t(t) * G + blockIdx.x] = tot[t];
}

// Fused local update and re-measure: W_r <- W_r - sum_{i<n} H[i][r] V_i
// (FP64 accumulation, skipped unless *gate == want when gate is given), then
// the dots of rows 0 .. nd-1 of Vd against the updated W_r (no weights), as
// bdots_kernel(Vd, nd, W) would compute them.  Vd may overlap W (the panel's
// own Gram): each thread reads back only the pixels it has just written, so
// V, Vd and W carry no __restrict__.
// The coefficients come through uniform (scalar) loads.  MODE 0: the dots
// of nd <= 2 B rows of Vd (the two local panels); MODE 1 (SELF): the
// panel's own Gram; MODE 2: both, the Gram appended as rows nd .. nd+B-1.
// In MODE 1 a skipped update (gate given and *gate != want) leaves the panel
// as it is and returns the Gram the caller already has (skip_gram, B x B)
// through the partials, and clears skip_zero[0 .. nzero).
template <int B, int NV, int PX, int MODE>
__global__ __launch_bounds__(256) void bupdate_dots_kernel(const float* V, size_t ld, int n,
                                                           const double* __restrict__ Hm, float* Wm, size_t ldw,
                                                           size_t Np, const int* __restrict__ gate, int want,
                                                           const float* Vd, int nd, double* __restrict__ partial,
                                                           const double* skip_gram, double* skip_zero, int nzero) {
  constexpr bool SELF = MODE == 1;
  constexpr int kRows = SELF ? B : (MODE == 2 ? 3 * B : 2 * B);
  __shared__ double red[4][DotsAcc<B, NV, PX>::K];
  __shared__ double tot[kRows * B];
  const bool upd = n > 0 && !(gate && *gate != want);  // uniform
  const int tid = threadIdx.x;
  const int G = gridDim.x;
  if (SELF) nd = 0;
  const int nt = (nd + (MODE == 0 ? 0 : B)) * B;  // entries written
  if (SELF && gate && !upd) {
    for (int t = tid; t < nt; t += 256) partial[size_t(t) * G + blockIdx.x] = blockIdx.x == 0 ? skip_gram[t] : 0.0;
    if (blockIdx.x == 0)
      for (int t = tid; t < nzero; t += 256) skip_zero[t] = 0.0;
    return;
  }
  for (int t = tid; t < nt; t += 256) tot[t] = 0.0;
  __syncthreads();
  const size_t nchunk = (Np + 256 * PX - 1) / (256 * PX);
  for (size_t ch = blockIdx.x; ch < nchunk; ch += G) {
    size_t pp[PX];
    float w[PX][B];
    double dd[PX];
#pragma unroll
    for (int k = 0; k < PX; k++) {
      const size_t p = ch * (256 * PX) + size_t(k) * 256 + tid;
      const bool in = p < Np;
      pp[k] = in ? p : Np - 1;
      dd[k] = in ? 1.0 : 0.0;
    }
    // the update in groups of PH pixels (FP64 accumulators of one group
    // live at a time); basis vectors outermost within a group, so the
    // group's loads of one vector are in flight together (per element the
    // same FMA order as bupdate_kernel)
    constexpr int PH = PX < 4 ? PX : 4;
#pragma unroll
    for (int k0 = 0; k0 < PX; k0 += PH) {
      double acc[PH][B];
#pragma unroll
      for (int k = 0; k < PH; k++)
#pragma unroll
        for (int r = 0; r < B; r++) acc[k][r] = Wm[size_t(r) * ldw + pp[k0 + k]];
      if (upd) {
        for (int i = 0; i < n; i++) {
          double h[B];
#pragma unroll
          for (int r = 0; r < B; r++) h[r] = Hm[i * B + r];
#pragma unroll
          for (int k = 0; k < PH; k++) {
            const double v = V[size_t(i) * ld + pp[k0 + k]];
#pragma unroll
            for (int r = 0; r < B; r++) acc[k][r] = fma(-h[r], v, acc[k][r]);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < PH; k++) {
#pragma unroll
        for (int r = 0; r < B; r++) w[k0 + k][r] = float(acc[k][r]);
        if (upd && dd[k0 + k] != 0.0) {
#pragma unroll
          for (int r = 0; r < B; r++) Wm[size_t(r) * ldw + pp[k0 + k]] = w[k0 + k][r];
        }
      }
    }
    if (MODE != 1) dots_chunk<B, NV, PX, false>(Vd, ld, nd, pp, w, dd, red, tot, tid);
    if (MODE != 0) dots_chunk<B, NV, PX, true>(Vd, ld, B, pp, w, dd, red, tot + nd * B, tid);
  }
  for (int t = tid; t < nt; t += 256) partial[size_t(t) * G + blockIdx.x] = tot[t];
}

// Second local pass needed?  DGKS-style test per column on the panel's
// squared norms before the first pass (measured Gram, rows n .. n+B-1 of H1)
// and after it (appended Gram, rows n .. n+B-1 of H2): the pass runs when a
// column's squared norm fell below kBDgksEta2 of what it was.
__global__ void bdgks_kernel(const double* __restrict__ H1, const double* __restrict__ H2, int n, int B,
                             int* __restrict__ flag, double* __restrict__ acc) {
  if (threadIdx.x != 0) return;
  bool need = false;
  for (int r = 0; r < B; r++) {
    const double before = H1[size_t(n + r) * B + r], after = H2[size_t(n + r) * B + r];
    need = need || !(after >= kBDgksEta2 * before);
  }
  *flag = need ? 1 : 0;
  if (need) acc[0] += 1.0;
}

// Fused CholQR step: X_r <- sum_{s<=r} X_s Rinv[s][r] in place (as
// btransform_kernel), then the Gram X^T X of the result (as
// bdots_kernel(X, B, X)), partial[(i * B + r) * G + g].
template <int B, int NV, int PX>
__global__ __launch_bounds__(256) void btransform_gram_kernel(float* __restrict__ X, size_t ld,
                                                              const double* __restrict__ Rinv, size_t Np,
                                                              double* __restrict__ partial) {
  __shared__ double red[4][DotsAcc<B, NV, PX>::K];
  __shared__ double tot[B * B];
  __shared__ double rs[B * B];
  const int tid = threadIdx.x;
  const int G = gridDim.x;
  for (int t = tid; t < B * B; t += 256) {
    tot[t] = 0.0;
    rs[t] = Rinv[t];
  }
  __syncthreads();
  const size_t nchunk = (Np + 256 * PX - 1) / (256 * PX);
  for (size_t ch = blockIdx.x; ch < nchunk; ch += G) {
    size_t pp[PX];
    float w[PX][B];
    double dd[PX];
#pragma unroll
    for (int k = 0; k < PX; k++) {
      const size_t p = ch * (256 * PX) + size_t(k) * 256 + tid;
      const bool in = p < Np;
      pp[k] = in ? p : Np - 1;
      dd[k] = in ? 1.0 : 0.0;
      double x[B];
#pragma unroll
      for (int r = 0; r < B; r++) x[r] = X[size_t(r) * ld + pp[k]];
#pragma unroll
      for (int r = 0; r < B; r++) {
        double a = 0;
#pragma unroll
        for (int q = 0; q <= r; q++) a += x[q] * rs[q * B + r];
        w[k][r] = float(a);
      }
      if (in) {
#pragma unroll
        for (int r = 0; r < B; r++) X[size_t(r) * ld + pp[k]] = w[k][r];
      }
    }
    dots_chunk<B, NV, PX, true>(X, ld, B, pp, w, dd, red, tot, tid);
  }
  for (int t = tid; t < B * B; t += 256) partial[size_t(t) * G + blockIdx.x] = tot[t];
}

// W_r <- W_r - sum_i H[i][r] V_i (FP64 accumulation); PX pixels per thread
// share each coefficient read from LDS.
template <int B, int PX>
__global__ __launch_bounds__(256) void bupdate_kernel(const float* __restrict__ V, size_t ld, int n,
                                                      const double* __restrict__ Hm, float* __restrict__ Wm, size_t ldw,
                                                      size_t Np, const int* __restrict__ gate, int want) {
  if (gate && *gate != want) return;  // uniform
  __shared__ double hs[kBMaxBasis * B];
  for (int t = threadIdx.x; t < n * B; t += blockDim.x) hs[t] = Hm[t];
  __syncthreads();
  const size_t step = size_t(gridDim.x) * 256 * PX;
  for (size_t p0 = size_t(blockIdx.x) * 256 * PX + threadIdx.x; p0 < Np; p0 += step) {
    size_t pp[PX];
    double acc[PX][B];
#pragma unroll
    for (int k = 0; k < PX; k++) {
      const size_t p = p0 + size_t(k) * 256;
      pp[k] = p < Np ? p : Np - 1;
#pragma unroll
      for (int r = 0; r < B; r++) acc[k][r] = Wm[size_t(r) * ldw + pp[k]];
    }
    for (int i = 0; i < n; i++) {
      double h[B];
#pragma unroll
      for (int r = 0; r < B; r++) h[r] = hs[i * B + r];
#pragma unroll
      for (int k = 0; k < PX; k++) {
        const double v = V[size_t(i) * ld + pp[k]];
#pragma unroll
        for (int r = 0; r < B; r++) acc[k][r] = fma(-h[r], v, acc[k][r]);
      }
    }
#pragma unroll
    for (int k = 0; k < PX; k++) {
      if (p0 + size_t(k) * 256 < Np) {
#pragma unroll
        for (int r = 0; r < B; r++) Wm[size_t(r) * ldw + pp[k]] = float(acc[k][r]);
      }
    }
  }
}

// partial[i][g] = sum over workgroup g's pixels of D[p] X_i[p]^2, i < n.
__global__ __launch_bounds__(256) void bwnorm_kernel(const float* __restrict__ X, size_t ld, int n,
                                                     const float* __restrict__ D, size_t Np,
                                                     double* __restrict__ partial) {
  __shared__ double sh[4];
  const int i = blockIdx.y;
  const float* x = X + size_t(i) * ld;
  double a = 0;
  for (size_t p = size_t(blockIdx.x) * 256 + threadIdx.x; p < Np; p += size_t(gridDim.x) * 256) {
    const double v = x[p];
    a = fma(double(D[p]) * v, v, a);
  }
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) partial[size_t(i) * gridDim.x + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// In place X_r <- sum_s X_s Rinv[s][r] (upper triangular Rinv), FP64.
template <int B>
__global__ __launch_bounds__(256) void btransform_kernel(float* __restrict__ X, size_t ld, const double* __restrict__ Rinv,
                                                         size_t Np, const int* __restrict__ skip = nullptr) {
  if (skip && *skip) return;  // uniform
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
//
// skip2 (optional): pass 1 sets *skip2 = 1 when the column-scaled factor
// R D^-1 (D = diag R) has Frobenius condition number below kBChol2Cond (an
// upper bound of its 2-norm condition number) and no pivot was floored;
// pass 2 then only sets Rinv = I and Rtot = R1, and the transform that
// follows is skipped.  With the Gram summed in FP64 one CholQR pass loses
// orthogonality only at eps_64 * cond^2, and the FP32 rounding of the result
// is not amplified by the condition number.
template <int B>
__global__ void bchol_kernel(const double* __restrict__ G, double* __restrict__ Rinv, double* __restrict__ R1,
                             double* __restrict__ Rtot, int pass, int* __restrict__ brk, int* __restrict__ skip2) {
  if (threadIdx.x != 0) return;
  if (pass == 2 && skip2 && *skip2) {
    for (int t = 0; t < B * B; t++) {
      Rinv[t] = (t / B == t % B) ? 1.0 : 0.0;
      Rtot[t] = R1[t];
    }
    return;
  }
  double R[B * B], Ri[B * B];
  bool floored = false;
  double gmax = 0;
  for (int r = 0; r < B; r++) gmax = fmax(gmax, G[r * B + r]);
  for (int t = 0; t < B * B; t++) R[t] = Ri[t] = 0;
  for (int j = 0; j < B; j++) {
    double d = G[j * B + j];
    for (int k = 0; k < j; k++) d -= R[k * B + j] * R[k * B + j];
    if (!(d > 1e-12 * gmax)) {
      d = fmax(1e-12 * gmax, 1e-300);
      atomicAdd(brk, 1);
      floored = true;
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
    if (skip2) {
      double a = 0, b = 0;
      for (int i = 0; i < B; i++)
        for (int j = i; j < B; j++) {
          const double x = R[i * B + j] / R[j * B + j], y = Ri[i * B + j] * R[i * B + i];
          a += x * x;
          b += y * y;
        }
      *skip2 = (!floored && sqrt(a * b) < kBChol2Cond) ? 1 : 0;
    }
  } else {
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        double a = 0;
        for (int k = 0; k < B; k++) a += R[i * B + k] * R1[k * B + j];
        Rtot[i * B + j] = a;
      }
  }
}

// T[i][k + r] += H1[i][r] + H2[i][r] for lo <= i < n, then mirrored: rows
// above the panel copy their column entries, and the panel's diagonal block is
// made symmetric by averaging.  One block; the phases are separated by
// barriers so that no entry is read while another thread updates it.
__global__ void btcol_kernel(double* __restrict__ T, int M, int k, int n, int B, int lo, const double* __restrict__ H1,
                             const double* __restrict__ H2) {
  for (int t = lo * B + threadIdx.x; t < n * B; t += blockDim.x) {
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

// Semi-orthogonal block steps.  A block step's first pass measures the new
// panel against the whole basis (H = V^T Pn, rows n .. n+B-1 holding the
// panel's own Gram); the update then removes only the panel's two preceding
// blocks [lo, n) -- the block three-term recurrence -- unless some overlap
// with an older vector, relative to the norm the local update leaves
// (||Pn_r||^2 - sum_{lo <= i < n} H[i][r]^2), exceeds kBSemiOrthTol, in which
// case the whole basis is removed (*flag = 1).  The level is kBSemiOrthTol
// unless set for verification (gpb_set_reorth_threshold).  A second Gram-Schmidt pass
// against the two blocks follows either way; T keeps the block-tridiagonal
// entries only.  acc[0] counts the whole-basis updates.
constexpr double kBSemiOrthTol = 2.44140625e-04;  // sqrt of the FP32 unit roundoff 2^-24

__global__ __launch_bounds__(256) void bgate_kernel(const double* __restrict__ H, int n, int lo, int B,
                                                    int* __restrict__ flag, double* __restrict__ acc, double thr) {
  __shared__ double est[kBMax];
  __shared__ double mx[4];
  const int t = threadIdx.x;
  if (t < B) {
    double e = H[size_t(n + t) * B + t];
    for (int i = lo; i < n; i++) e -= H[size_t(i) * B + t] * H[size_t(i) * B + t];
    est[t] = fmax(e, 1e-300);
  }
  __syncthreads();
  double m = 0;
  for (int q = t; q < lo * B; q += 256) m = fmax(m, fabs(H[q]) / sqrt(est[q % B]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((t & 63) == 0) mx[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    const bool f = fmax(fmax(mx[0], mx[1]), fmax(mx[2], mx[3])) > thr;
    *flag = f ? 1 : 0;
    if (f) acc[0] += 1.0;
  }
}

// Estimated semi-orthogonality (the block form of the omega recurrence of
// Simon's partial reorthogonalisation, as in Grimes, Lewis and Simon's
// shifted block Lanczos), in place of measuring each new panel against the
// whole basis.  Om[l][c] estimates V_l . P_c for the panel P in question.
// From A P_k = sum_{l<n} V_l T[l][P_k] + P_n R (R = Rtot of the new panel),
// for an old row r < lo (below the two local panels [lo, n)):
//   (V_r . P_n) R = sum_{l<n} T[r][l] Om_k[l] - sum_{lo<=i<n} Om[r][i] H[i]
// with Om_k rows k .. n-1 the identity and rows lo .. k-1 (the previous
// panel, orthogonalised against P_k) at rounding level, which the sum leaves
// out, H = the measured local coefficients (rows lo .. n-1 of this step's
// first pass), Om[r][i] = Om_l (i < k) or Om_k (i >= k).  X = the right-hand
// side plus a rounding term eps (sum_{l<n} |T[r][l]| + B max |H|) of the sign
// of each entry (the single-vector recurrence's noise model).  The gate then compares |X[r][c]|
// with the norm the local update leaves, as bgate_kernel does with measured
// overlaps; the step after a flagged one is flagged too (pending).
constexpr double kBOmEps = 5.9604644775390625e-08;  // 2^-24, unit roundoff of the FP32 basis
__global__ __launch_bounds__(256) void bomega_est_kernel(const double* __restrict__ T, int M, int k, int n, int lo,
                                                         int B, const double* __restrict__ H,
                                                         const double* __restrict__ OmK,
                                                         const double* __restrict__ OmL, double* __restrict__ X,
                                                         int* __restrict__ flag, int* __restrict__ pending,
                                                         double* __restrict__ acc, double thr) {
  __shared__ double est[kBMax];
  __shared__ double mx[4];
  const int t = threadIdx.x;
  if (t < B) {
    double e = H[size_t(n + t) * B + t];
    for (int i = lo; i < n; i++) e -= H[size_t(i) * B + t] * H[size_t(i) * B + t];
    est[t] = fmax(e, 1e-300);
  }
  // |H|_max over the local rows, for the rounding term
  double a = 0;
  for (int q = lo * B + t; q < n * B; q += 256) a = fmax(a, fabs(H[q]));
  for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
  if ((t & 63) == 0) mx[t >> 6] = a;
  __syncthreads();
  const double hn = fmax(fmax(mx[0], mx[1]), fmax(mx[2], mx[3]));
  __syncthreads();
  double m = 0;
  for (int q = t; q < lo * B; q += 256) {
    const int r = q / B, c = q - r * B;
    const double* Tr = T + size_t(r) * M;
    double x = 0, tn = 0;
    // sum over l in order, the loads of 16 terms issued before their FMAs
    for (int l0 = 0; l0 < lo; l0 += 16) {
      double tv[16], ov[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        tv[u] = l0 + u < lo ? Tr[l0 + u] : 0.0;
        ov[u] = l0 + u < lo ? OmK[size_t(l0 + u) * B + c] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; u++)
        if (l0 + u < lo) {
          x += tv[u] * ov[u];
          tn += fabs(tv[u]);
        }
    }
    x += Tr[k + c];  // rows k .. n-1 of Om_k: the identity
    // the two local panels, eight terms' loads at a time
    for (int i0 = lo; i0 < n; i0 += 8) {
      double ov[8], hv[8], tv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int i = i0 + u;
        const bool in = i < n;
        ov[u] = in ? (i < k ? OmL[size_t(r) * B + (i - lo)] : OmK[size_t(r) * B + (i - k)]) : 0.0;
        hv[u] = in ? H[size_t(i) * B + c] : 0.0;
        tv[u] = in ? Tr[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (i0 + u < n) {
          x -= ov[u] * hv[u];
          tn += fabs(tv[u]);
        }
    }
    const double noise = kBOmEps * (tn + B * hn);
    x += x >= 0 ? noise : -noise;
    X[q] = x;
    m = fmax(m, fabs(x) / sqrt(est[c]));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((t & 63) == 0) mx[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    const bool viol = fmax(fmax(mx[0], mx[1]), fmax(mx[2], mx[3])) > thr;
    const int pend = *pending;
    const bool f = viol || pend;
    *flag = f ? 1 : 0;
    *pending = (viol && !pend) ? 1 : 0;
    if (f) acc[0] += 1.0;
  }
}

// After CholQR of the new panel: Om_new[r] = X[r] R^-1 (R upper triangular)
// for the old rows r < lo, rounding level for the rows lo .. n-1 the step
// orthogonalised against; rounding level for every row when the
// whole-basis pass ran (flag, or flag == nullptr after a full
// Gram-Schmidt step).
__global__ __launch_bounds__(256) void bomega_finish_kernel(const double* __restrict__ X, const double* __restrict__ R,
                                                            double* __restrict__ Om, int lo, int n, int B,
                                                            const int* __restrict__ flag) {
  const bool whole = !flag || *flag;
  for (int r = threadIdx.x; r < n; r += 256) {
    double* o = Om + size_t(r) * B;
    if (whole || r >= lo) {
      for (int c = 0; c < B; c++) o[c] = kBOmEps;
      continue;
    }
    const double* x = X + size_t(r) * B;
    for (int c = 0; c < B; c++) {
      double v = x[c];
      for (int a = 0; a < c; a++) v -= o[a] * R[size_t(a) * B + c];
      o[c] = v / R[size_t(c) * B + c];
    }
  }
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

// Launch shapes of the panel kernels per block size: NV basis vectors per
// dots group, PX pixels per thread in the dots and update kernels.
template <int B>
struct BlockShape {
  // (B = 8: NV 2, PX 8 measured fastest of (4, 4), (2, 8), (2, 16) -- 19 %
  // less dots time than (4, 4) at 4096^2; (2, 16) runs at occupancy 1)
  static constexpr int NV = B <= 4 ? 8 : (B < 8 ? 4 : 2);
  static constexpr int PXD = B < 8 ? 4 : (B == 8 ? 8 : 2);
  static constexpr int PXU = B <= 8 ? 2 : 1;
};
constexpr int kBDotsGrid = 1024;

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
    NC = int(std::min<size_t>(kBDotsGrid, div_up(Np, size_t(256) * BlockShape<B>::PXD)));
    partial = c.ws<double>("beig.partial", size_t(kBMaxBasis) * kBMax * kBDotsGrid);
    int dx[kStencilHalf], dy[kStencilHalf];
    stencil_upper_offsets(dx, dy);
    for (int j = 0; j < kStencilHalf; j++) {
      A.o[j] = st.off_host[j];
      A.lo[j] = dy[j] * kBTileW + dx[j];
    }
    vg = std::min<unsigned>(2048, div_up(Np, 256));
    ug = std::min<unsigned>(2048, div_up(Np, size_t(256) * BlockShape<B>::PXU));
  }
  unsigned ug;
  void spmv(const float* X, float* Y) {
    if (band == 0) band = tune_band(X, Y);
    KScope ks(c, kKSpmv, double(st.W) * st.H * (4.0 * kStencilHalf + 4 + 8.0 * B));
    launch_spmv_band(X, Y, band);
  }
  int band = 0;
  void launch_spmv_band(const float* X, float* Y, int bnd) {
    const int ntiles = int(div_up(st.W, kBTX) * div_up(st.H, kBTY));
    const dim3 grid(unsigned(div_up(ntiles, kNumXcd) * kNumXcd)), block(kBTX, kBTY);
    if (c.spmv_fp64)
      hipLaunchKernelGGL((bspmv_kernel<B, double>), grid, block, 0, c.stream, st.wup, st.s, X, Np, Y, Np, st.W, st.H,
                         st.Wp, st.Hp, Np, A, bnd);
    else
      hipLaunchKernelGGL((bspmv_kernel<B, float>), grid, block, 0, c.stream, st.wup, st.s, X, Np, Y, Np, st.W, st.H,
                         st.Wp, st.Hp, Np, A, bnd);
    GPB_LAUNCH_CHECK();
  }
  // Band width of the block SpMV's tile order for this shape and B: timed
  // once per context (2 applications after a warm-up per candidate) on the
  // solve's own operands, which the calls overwrite with the same values.
  int tune_band(const float* X, float* Y) {
    const int tiles_x = int(div_up(st.W, kBTX));
    const uint64_t key = ((uint64_t(st.W) << 32) | uint32_t(st.H)) ^ (uint64_t(B) << 52) ^ (1ull << 61) ^
                         (c.spmv_fp64 ? (1ull << 62) : 0);
    auto it = c.spmv_band.find(key);
    if (it != c.spmv_band.end()) return it->second;
    hipEvent_t e0, e1;
    GPB_HIP(hipEventCreate(&e0));
    GPB_HIP(hipEventCreate(&e1));
    int best = tiles_x;
    float best_ms = 1e30f;
    for (int bnd : {2, 4, 8, tiles_x}) {
      if (bnd > tiles_x) continue;
      launch_spmv_band(X, Y, bnd);
      GPB_HIP(hipEventRecord(e0, c.stream));
      for (int r = 0; r < 2; r++) launch_spmv_band(X, Y, bnd);
      GPB_HIP(hipEventRecord(e1, c.stream));
      GPB_HIP(hipEventSynchronize(e1));
      float ms = 0;
      GPB_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (c.verbose)
        std::fprintf(stderr, "gpb: block SpMV tuning %dx%d B=%d: band %d: %.1f us\n", st.W, st.H, B, bnd,
                     ms * 1e3f / 2);
      if (ms < best_ms) {
        best_ms = ms;
        best = bnd;
      }
    }
    GPB_HIP(hipEventDestroy(e0));
    GPB_HIP(hipEventDestroy(e1));
    c.spmv_band[key] = best;
    return best;
  }
  // dst[i * B + r] = V_i . (D) W_r, i < n
  void dots(const float* V, int n, const float* Wm, const float* D, double* dst) {
    {
      KScope ks(c, kKDots, double(Np) * (4.0 * n + 4.0 * B + (D ? 4 : 0)));
      if (n > kBMaxBasis) throw ArgError("block Lanczos: basis above the dots kernel's capacity");
      hipLaunchKernelGGL((bdots_kernel<B, BlockShape<B>::NV, BlockShape<B>::PXD>), dim3(NC), dim3(256), 0, c.stream, V,
                         Np, n, Wm, Np, D, Np, partial, nullptr);
      GPB_LAUNCH_CHECK();
    }
    launch_reduce(c, partial, NC, n * B, dst);
  }
  // dots(V, n, Wm, nullptr, dst) only when *gate != 0 (dst untouched otherwise)
  void dots_gated(const float* V, int n, const float* Wm, double* dst, const int* gate) {
    if (n <= 0) return;
    {
      KScope ks(c, kKDots, 0.0);
      if (n > kBMaxBasis) throw ArgError("block Lanczos: basis above the dots kernel's capacity");
      hipLaunchKernelGGL((bdots_kernel<B, BlockShape<B>::NV, BlockShape<B>::PXD>), dim3(NC), dim3(256), 0, c.stream, V,
                         Np, n, Wm, Np, nullptr, Np, partial, gate);
      GPB_LAUNCH_CHECK();
    }
    launch_reduce_gated(c, partial, NC, n * B, dst, gate);
  }
  // W_r -= sum_i Hm[i][r] V_i, i < n; with gate, only when *gate == want
  void update(const float* V, int n, const double* Hm, float* Wm, const int* gate = nullptr, int want = 1) {
    KScope ks(c, kKUpdate, gate ? 0.0 : double(Np) * (4.0 * n + 8.0 * B));
    hipLaunchKernelGGL((bupdate_kernel<B, BlockShape<B>::PXU>), dim3(ug), dim3(256), 0, c.stream, V, Np, n, Hm, Wm, Np,
                       Np, gate, want);
    GPB_LAUNCH_CHECK();
  }
  // out_c = sum_i Y[i][c] V_i, c < nk
  void ritz(const float* V, int m, int nk, const double* Ydev, float* out) { launch_ritz(c, V, Np, m, nk, Ydev, out, Np); }
  // dst[i] = ||D^1/2 X_i||^2, i < n
  void wnorms(const float* X, int n, double* dst) {
    {
      KScope ks(c, kKDots, double(Np) * (4.0 * n + 4));
      hipLaunchKernelGGL(bwnorm_kernel, dim3(NC, n), dim3(256), 0, c.stream, X, Np, n, st.D, Np, partial);
      GPB_LAUNCH_CHECK();
    }
    launch_reduce(c, partial, NC, n, dst);
  }
  // W_r -= sum_{i<n} Hm[i][r] V_i (only when *gate == want, with gate), then
  // dst[i * B + r] = Vd_i . W_r for i < nd: update and dots(Vd, nd, W) in one
  // pass over the pixels, with the same sums as the two launches.
  // With append, dst rows nd .. nd+B-1 receive W^T W of the updated panel.
  void update_dots(const float* V, int n, const double* Hm, float* Wm, const int* gate, int want, const float* Vd,
                   int nd, double* dst, bool append = false) {
    if (nd > 2 * B || n > kBMaxBasis) throw ArgError("block Lanczos: fused pass takes at most two local panels");
    {
      KScope ks(c, kKUpdate, double(Np) * (4.0 * n + 8.0 * B));
      if (append)
        hipLaunchKernelGGL((bupdate_dots_kernel<B, BlockShape<B>::NV, BlockShape<B>::PXD, 2>), dim3(NC), dim3(256), 0,
                           c.stream, V, Np, n, Hm, Wm, Np, Np, gate, want, Vd, nd, partial, nullptr, nullptr, 0);
      else
        hipLaunchKernelGGL((bupdate_dots_kernel<B, BlockShape<B>::NV, BlockShape<B>::PXD, 0>), dim3(NC), dim3(256), 0,
                           c.stream, V, Np, n, Hm, Wm, Np, Np, gate, want, Vd, nd, partial, nullptr, nullptr, 0);
      GPB_LAUNCH_CHECK();
    }
    launch_reduce(c, partial, NC, (nd + (append ? B : 0)) * B, dst);
  }
  // W_r -= sum_{i<n} Hm[i][r] V_i, then G = W^T W (B x B) of the result.
  // With gate: only when *gate == 1; otherwise G = skip_gram (the Gram of the
  // unchanged panel) and Hm[0 .. n*B) is cleared (no coefficients applied).
  void update_gram(const float* V, int n, const double* Hm, float* Wm, double* G, const int* gate = nullptr,
                   const double* skip_gram = nullptr) {
    if (n > kBMaxBasis) throw ArgError("block Lanczos: basis above the dots kernel's capacity");
    {
      KScope ks(c, kKUpdate, gate ? 0.0 : double(Np) * (4.0 * n + 8.0 * B));
      hipLaunchKernelGGL((bupdate_dots_kernel<B, BlockShape<B>::NV, BlockShape<B>::PXD, 1>), dim3(NC), dim3(256), 0,
                         c.stream, V, Np, n, Hm, Wm, Np, Np, gate, 1, Wm, 0, partial, skip_gram,
                         const_cast<double*>(Hm), n * B);
      GPB_LAUNCH_CHECK();
    }
    launch_reduce(c, partial, NC, B * B, G);
  }
  // CholQR2 of the panel X (in place); Rtot = R of X_in = X_out Rtot.  With
  // gram_ready, G already holds X^T X (from update_gram).  The fused form
  // computes the second pass' Gram in the first pass' transform; with dgks
  // it also skips the second pass for well-conditioned panels (bchol_kernel).
  void cholqr2(float* X, double* G, double* Rinv, double* R1, double* Rtot, int* brk, bool gram_ready = false) {
    if (fused) {
      int* skip2 = dgks ? c.ws<int>("beig.skip2", 1) : nullptr;
      if (!gram_ready) dots(X, B, X, nullptr, G);
      hipLaunchKernelGGL(bchol_kernel<B>, dim3(1), dim3(64), 0, c.stream, G, Rinv, R1, Rtot, 1, brk, skip2);
      GPB_LAUNCH_CHECK();
      {
        KScope ks(c, kKUpdate, double(Np) * 8.0 * B);
        hipLaunchKernelGGL((btransform_gram_kernel<B, BlockShape<B>::NV, BlockShape<B>::PXD>), dim3(NC), dim3(256), 0,
                           c.stream, X, Np, Rinv, Np, partial);
        GPB_LAUNCH_CHECK();
      }
      launch_reduce(c, partial, NC, B * B, G);
      hipLaunchKernelGGL(bchol_kernel<B>, dim3(1), dim3(64), 0, c.stream, G, Rinv, R1, Rtot, 2, brk, skip2);
      GPB_LAUNCH_CHECK();
      KScope ks(c, kKUpdate, skip2 ? 0.0 : double(Np) * 8.0 * B);
      hipLaunchKernelGGL(btransform_kernel<B>, dim3(vg), dim3(256), 0, c.stream, X, Np, Rinv, Np, skip2);
      GPB_LAUNCH_CHECK();
      return;
    }
    for (int pass = 1; pass <= 2; pass++) {
      if (pass > 1 || !gram_ready) dots(X, B, X, nullptr, G);
      hipLaunchKernelGGL(bchol_kernel<B>, dim3(1), dim3(64), 0, c.stream, G, Rinv, R1, Rtot, pass, brk, nullptr);
      GPB_LAUNCH_CHECK();
      KScope ks(c, kKUpdate, double(Np) * 8.0 * B);
      hipLaunchKernelGGL(btransform_kernel<B>, dim3(vg), dim3(256), 0, c.stream, X, Np, Rinv, Np);
      GPB_LAUNCH_CHECK();
    }
  }
  bool fused = true;
  bool dgks = true;  // gated second local pass and CholQR second pass (fused form only)
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
  ops.fused = c.block_fusion >= 1;
  ops.dgks = c.block_fusion >= 2;
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
  // panels are built in place and the SpMV writes interior pixels only: the
  // halo entries of every column must read zero
  launch_fill(c, V, size_t(M + B) * Np, 0.f);
  launch_fill(c, V2, size_t(M + B) * Np, 0.f);

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
  std::vector<double> th, Y;
  // restart exchanges through pinned host buffers (asynchronous DMA; read
  // after a stream synchronisation, upload buffers rewritten only after the
  // next one)
  double* Th = c.host_ws<double>("beig.h.T", size_t(M) * M);
  double* Rh = c.host_ws<double>("beig.h.R", size_t(B) * B);
  double* Gd = c.host_ws<double>("beig.h.G", size_t(B) * B);
  int* brk_h = c.host_ws<int>("beig.h.brk", 1);
  double* Yk = c.host_ws<double>("beig.h.Y", size_t(M) * keep);
  double* nh = c.host_ws<double>("beig.h.nrm", size_t(K));
  double* T0 = c.host_ws<double>("beig.h.T0", size_t(M) * M);
  std::vector<double> resid(K, 0.0);
  bool converged = false;
  int kcur = 0;
  // semi-orthogonal block steps for targets the FP64 refinement does not follow
  const bool semi = c.reorth_mode >= 1 && eig_partial_reorth(opt);
  int* gflag = c.ws<int>("beig.gflag", 1);
  double* gacc = c.ws<double>("beig.gacc", 1);
  int* dflag = c.ws<int>("beig.dflag", 1);
  double* dacc = c.ws<double>("beig.dacc", 1);
  // With reorthogonalisation mode 2 (the default) the semi-orthogonal steps
  // estimate the new panel's overlaps with the older basis by the block omega
  // recurrence (bomega_est_kernel) and measure them only when it triggers;
  // mode 1 measures them on every step.
  const bool om_est = semi && c.reorth_mode == 2;
  // the gate's level: sqrt of the FP32 unit roundoff unless set for verification
  const double semi_thr = c.pro_threshold > 0 ? c.pro_threshold : kBSemiOrthTol;
  double* om_buf = om_est ? c.ws<double>("beig.om", size_t(3) * kBMaxBasis * kBMax) : nullptr;
  double* om_x = om_est ? c.ws<double>("beig.omx", size_t(kBMaxBasis) * kBMax) : nullptr;
  int* gpend = om_est ? c.ws<int>("beig.gpend", 1) : nullptr;
  // om slots: the estimates for the previous panel, the current one, and the
  // one being built
  int om_l = 0, om_k = 1, om_n = 2;
  auto om_slot = [&](int i) { return om_buf + size_t(i) * kBMaxBasis * kBMax; };
  hipLaunchKernelGGL(fill_f64_kernel, dim3(1), dim3(1), 0, c.stream, dacc, size_t(1), 0.0);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(fill_f64_kernel, dim3(1), dim3(1), 0, c.stream, gacc, size_t(1), 0.0);
  GPB_LAUNCH_CHECK();
  // The residual panel's overlap with the kept Ritz vectors (inherited from
  // the basis' semi-orthogonality) would feed every local step of the next
  // cycle: Gram-Schmidt it against them twice, re-orthonormalise.
  auto clean_panel = [&](float* Vn) {
    float* P = Vn + size_t(keep) * Np;
    for (int pass = 0; pass < 2; pass++) {
      ops.dots(Vn, keep, P, nullptr, H1);
      ops.update(Vn, keep, H1, P);
    }
    ops.cholqr2(P, G, Rinv, R1, Rtot, brk);
  };
  // The clean-up depends only on the new basis: it is queued before the host
  // reads the Ritz norms and runs while the host decides (after the last
  // restart it is work on a panel no output uses).  Not under per-kernel
  // timing, which reads its events at the restart's first sync.
  const bool early_clean = semi && !c.kt.on;
  hipEvent_t norms_ready = nullptr;
  if (early_clean) GPB_HIP(hipEventCreateWithFlags(&norms_ready, hipEventDisableTiming));
  struct EventGuard {
    hipEvent_t e;
    ~EventGuard() {
      if (e) (void)hipEventDestroy(e);
    }
  } norms_guard{norms_ready};
  while (true) {
    const int cycle_start = k;
    if (om_est) {
      // the cycle's first panel was cleaned against the kept vectors
      hipLaunchKernelGGL(bomega_finish_kernel, dim3(1), dim3(256), 0, c.stream, om_x, Rtot, om_slot(om_k), 0, k, B,
                         nullptr);
      GPB_LAUNCH_CHECK();
      hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(1), 0, c.stream, gpend, 0);
      GPB_LAUNCH_CHECK();
    }
    while (k + B <= M) {
      float* Pk = V + size_t(k) * Np;
      const int n = k + B;
      float* Pn = V + size_t(n) * Np;  // the next panel is built in place
      ops.spmv(Pk, Pn);
      matvecs += B;
      if (!semi || k == cycle_start) {
        // whole-basis classical Gram-Schmidt twice (the first step of a
        // cycle couples to every kept Ritz vector)
        ops.dots(V, n, Pn, nullptr, H1);
        ops.update(V, n, H1, Pn);
        ops.dots(V, n, Pn, nullptr, H2);
        if (ops.fused) {
          ops.update_gram(V, n, H2, Pn, G);  // the second pass and CholQR's first Gram
        } else {
          ops.update(V, n, H2, Pn);
        }
        hipLaunchKernelGGL(btcol_kernel, dim3(1), dim3(256), 0, c.stream, T, M, k, n, B, 0, H1, H2);
        GPB_LAUNCH_CHECK();
      } else {
        const int lo = k - B;
        if (om_est) {
          // measure the two local panels and the panel's own Gram (rows lo ..
          // n+B-1), estimate the rest, measure it only when the estimate
          // triggers the whole-basis pass
          ops.dots(V + size_t(lo) * Np, n + B - lo, Pn, nullptr, H1 + size_t(lo) * B);
          hipLaunchKernelGGL(bomega_est_kernel, dim3(1), dim3(256), 0, c.stream, T, M, k, n, lo, B, H1, om_slot(om_k),
                             om_slot(om_l), om_x, gflag, gpend, gacc, semi_thr);
          GPB_LAUNCH_CHECK();
          ops.dots_gated(V, lo, Pn, H1, gflag);
        } else {
          ops.dots(V, n + B, Pn, nullptr, H1);  // rows n .. n+B-1: the panel's own Gram
          hipLaunchKernelGGL(bgate_kernel, dim3(1), dim3(256), 0, c.stream, H1, n, lo, B, gflag, gacc, semi_thr);
          GPB_LAUNCH_CHECK();
        }
        ops.update(V, n, H1, Pn, gflag, 1);
        if (ops.fused && ops.dgks) {
          // local update with the re-measure of the two local panels and
          // the panel's new Gram; the second local pass only when the DGKS
          // test asks for it (else its coefficients are cleared, and the
          // appended Gram is CholQR's first)
          ops.update_dots(V + size_t(lo) * Np, n - lo, H1 + size_t(lo) * B, Pn, gflag, 0, V + size_t(lo) * Np, n - lo,
                          H2 + size_t(lo) * B, true);
          hipLaunchKernelGGL(bdgks_kernel, dim3(1), dim3(64), 0, c.stream, H1, H2, n, B, dflag, dacc);
          GPB_LAUNCH_CHECK();
          ops.update_gram(V + size_t(lo) * Np, n - lo, H2 + size_t(lo) * B, Pn, G, dflag, H2 + size_t(n) * B);
        } else if (ops.fused) {
          // local update with the re-measure of the two local panels, then
          // the second local pass with CholQR's first Gram
          ops.update_dots(V + size_t(lo) * Np, n - lo, H1 + size_t(lo) * B, Pn, gflag, 0, V + size_t(lo) * Np, n - lo,
                          H2 + size_t(lo) * B);
          ops.update_gram(V + size_t(lo) * Np, n - lo, H2 + size_t(lo) * B, Pn, G);
        } else {
          ops.update(V + size_t(lo) * Np, n - lo, H1 + size_t(lo) * B, Pn, gflag, 0);
          ops.dots(V + size_t(lo) * Np, n - lo, Pn, nullptr, H2 + size_t(lo) * B);
          ops.update(V + size_t(lo) * Np, n - lo, H2 + size_t(lo) * B, Pn);
        }
        hipLaunchKernelGGL(btcol_kernel, dim3(1), dim3(256), 0, c.stream, T, M, k, n, B, lo, H1, H2);
        GPB_LAUNCH_CHECK();
      }
      ops.cholqr2(Pn, G, Rinv, R1, Rtot, brk, ops.fused);
      if (om_est) {
        const bool full = !semi || k == cycle_start;
        hipLaunchKernelGGL(bomega_finish_kernel, dim3(1), dim3(256), 0, c.stream, om_x, Rtot, om_slot(om_n),
                           full ? n : k - B, n, B, full ? nullptr : gflag);
        GPB_LAUNCH_CHECK();
        const int t = om_l;
        om_l = om_k;
        om_k = om_n;
        om_n = t;
      }
      k += B;
    }
    kcur = k;  // projected size
    // Rayleigh-Ritz on the host (smallest first)
    GPB_HIP(hipMemcpyAsync(Th, T, sizeof(double) * M * M, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipMemcpyAsync(Rh, Rtot, sizeof(double) * B * B, hipMemcpyDeviceToHost, c.stream));
    // weighted Gram of the residual panel
    ops.dots(V + size_t(kcur) * Np, B, V + size_t(kcur) * Np, st.D, G);
    *brk_h = 0;
    GPB_HIP(hipMemcpyAsync(Gd, G, sizeof(double) * B * B, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipMemcpyAsync(brk_h, brk, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    c.kt.fold();
    // A rank-deficient panel means the block Krylov space became (nearly)
    // invariant; the caller falls back to the single-vector solver.
    if (*brk_h > 0) {
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
    for (int j = 0; j < kcur; j++)
      for (int i = 0; i < nk; i++) Yk[size_t(j) * nk + i] = Y[size_t(j) * kcur + i];
    GPB_HIP(hipMemcpyAsync(Ydev, Yk, sizeof(double) * kcur * nk, hipMemcpyHostToDevice, c.stream));
    ops.ritz(V, kcur, nk, Ydev, V2);
    GPB_HIP(hipMemcpyAsync(V2 + size_t(nk) * Np, V + size_t(kcur) * Np, size_t(B) * Np * sizeof(float),
                           hipMemcpyDeviceToDevice, c.stream));
    // ||D^1/2 u_i|| for the wanted Ritz vectors
    ops.wnorms(V2, K, nrm);
    GPB_HIP(hipMemcpyAsync(nh, nrm, sizeof(double) * K, hipMemcpyDeviceToHost, c.stream));
    if (early_clean) {
      GPB_HIP(hipEventRecord(norms_ready, c.stream));
      clean_panel(V2);
      GPB_HIP(hipEventSynchronize(norms_ready));
    } else {
      GPB_HIP(hipStreamSynchronize(c.stream));
    }
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
      const double un = nh[size_t(i)];
      resid[size_t(i)] = std::sqrt(std::max(rr, 0.0)) / std::sqrt(std::max(un, 1e-300));
      worst = std::max(worst, resid[size_t(i)]);
    }
    stats.max_residual = worst;
    if (c.verbose)
      std::fprintf(stderr, "gpb: block Lanczos %dx%d (block %d): %d operator applications, residual %.3e\n", st.W,
                   st.H, B, matvecs, worst);
    converged = conv_count(opt, th.data(), resid.data()) == K;
    if (converged || matvecs >= opt.max_matvecs) break;
    std::fill(T0, T0 + size_t(M) * M, 0.0);
    for (int i = 0; i < keep; i++) T0[size_t(i) * M + i] = th[size_t(i)];
    GPB_HIP(hipMemcpyAsync(T, T0, sizeof(double) * M * M, hipMemcpyHostToDevice, c.stream));
    k = keep;
    stats.restarts++;
    if (semi && !early_clean) clean_panel(V);
  }
  if (semi) {
    double g = 0, d = 0;
    GPB_HIP(hipMemcpyAsync(&g, gacc, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipMemcpyAsync(&d, dacc, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
    stats.second_passes = int(g);
    if (c.verbose && ops.dgks)
      std::fprintf(stderr, "gpb: block Lanczos: %d whole-basis passes, %d second local passes\n", int(g), int(d));
  }
  stats.iterations = matvecs;
  stats.converged = 0;
  stats.converged = th.empty() ? 0 : conv_count(opt, th.data(), resid.data());
  for (int i = 0; i < K; i++) {
    evals[i] = th[size_t(i)];
    launch_to_general(c, V + size_t(i) * Np, st.s, d_evecs + size_t(i) * Np, Np);
  }
  if (!converged)
    throw NoConvergence("block Lanczos stopped after " + std::to_string(matvecs) +
                        " operator applications with residual " + std::to_string(stats.max_residual));
}

}  // namespace

// partial[(2 i) G + g] = sum D (y_i - lam_i u_i)^2, partial[(2 i + 1) G + g] =
// sum D u_i^2 over workgroup g's pixels.
__global__ __launch_bounds__(256) void resid_parts_kernel(const float* __restrict__ U, const float* __restrict__ Y,
                                                          const float* __restrict__ D, const double* __restrict__ lam,
                                                          size_t Np, double* __restrict__ partial) {
  __shared__ double sh[2][4];
  const int i = blockIdx.y;
  const double l = lam[i];
  double a = 0, b = 0;
  for (size_t p = size_t(blockIdx.x) * 256 + threadIdx.x; p < Np; p += size_t(gridDim.x) * 256) {
    const double u = U[size_t(i) * Np + p], d = D[p];
    const double r = double(Y[size_t(i) * Np + p]) - l * u;
    a = fma(d * r, r, a);
    b = fma(d * u, u, b);
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[size_t(2 * i) * gridDim.x + blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    partial[size_t(2 * i + 1) * gridDim.x + blockIdx.x] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// Measured relative residuals ||(D - W) v_i - lam_i D v_i|| / ||D v_i|| of
// K <= 9 pairs (v generalised, padded layout): u = D^1/2 v, y = L u, and
// (D - W) v - lam D v = D^1/2 (y - lam u), D v = D^1/2 u.
double measured_max_residual(Context& c, const Stencil& st, const float* d_evecs, const double* evals, int K,
                             double* resid) {
  if (K < 1 || K > 9) throw ArgError("residual check takes 1 to 9 pairs");
  const size_t Np = st.Np;
  float* U = c.ws<float>("res.U", size_t(9) * Np);
  float* Y = c.ws<float>("res.Y", size_t(9) * Np);
  launch_fill(c, U, size_t(9) * Np, 0.f);
  const unsigned vg = std::min<unsigned>(2048, div_up(Np, 256));
  for (int i = 0; i < K; i++) {
    hipLaunchKernelGGL(to_uspace_kernel, dim3(vg), dim3(256), 0, c.stream, d_evecs + size_t(i) * Np, st.D,
                       U + size_t(i) * Np, Np);
    GPB_LAUNCH_CHECK();
  }
  launch_spmv(c, st, U, Y);
  if (K > 1) {
    BlockOps<8> ops(c, st);
    ops.spmv(U + Np, Y + Np);  // columns 1..8 (beyond K: zero vectors)
  }
  double* lam = c.ws<double>("res.lam", 9);
  double* part = c.ws<double>("res.part", size_t(18) * 1024);
  double* sums = c.ws<double>("res.sums", 18);
  GPB_HIP(hipMemcpyAsync(lam, evals, sizeof(double) * K, hipMemcpyHostToDevice, c.stream));
  const unsigned G = std::min<unsigned>(1024, div_up(Np, 256));
  hipLaunchKernelGGL(resid_parts_kernel, dim3(G, K), dim3(256), 0, c.stream, U, Y, st.D, lam, Np, part);
  GPB_LAUNCH_CHECK();
  launch_reduce(c, part, int(G), 2 * K, sums);
  std::vector<double> h(static_cast<size_t>(2 * K));
  GPB_HIP(hipMemcpyAsync(h.data(), sums, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
  double worst = 0;
  for (int i = 0; i < K; i++) {
    const double r = std::sqrt(std::max(h[size_t(2 * i)], 0.0)) / std::sqrt(std::max(h[size_t(2 * i + 1)], 1e-300));
    if (resid) resid[i] = r;
    worst = std::max(worst, r);
  }
  return worst;
}

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

// Stage (b): colour front end, half-disc chi-squared gradients, mPb combine
// and oriented nonmax.
//
// References: mex_pb_parts_final_selected.cpp:76-227 (gamma, Lab, quantise,
// 12 hist_gradient_2D calls), lib_image.cc:3120-3231 (the histogram
// gradient), multiscalePb.cpp:216-261 (weighted sum, nonmax, clip),
// nonmax_channels.cpp:8-80 and lib_image.cc:4083-4200 (oriented nonmax).
//
// The chi-squared kernel keeps one wavefront = one 8x8 pixel tile.  Each lane
// owns a private histogram in LDS laid out [slice/4][bin][lane] with four 8-bit
// slice counters per dword, so every increment is a conflict-free ds_add_u32
// whose address depends on the lane's own label only.  Disc offsets are walked
// in increasing radius, so the three radii of a channel are snapshots of one
// pass.  Per snapshot the 16 slice histograms are folded into the half-disc
// chi-squared of all 8 orientations using linearity of the bin smoothing:
// with a_s = 1 / (c . h_s) (c = column sums of the truncated smoothing
// matrix G), sum = G * sum_s a_s h_s and diff_o = G * V_o, where
// V_{o+1} = V_o + 2 a_o h_o - 2 a_{o+8} h_{o+8}.
#include <cstdio>
#include <array>
#include <mutex>

#include "gpb_pipeline.hpp"

namespace gpb {

// ---- colour front end ------------------------------------------------------
// Gamma 2.5, RGB->XYZ->Lab (D65), normalise and quantise to 25 bins in FP64
// (lib_image.cc:123-138, 172-185, 216-245, 143-167, 2315-2344).  The gamma
// curve is a 256-entry table of pow(c/255, 2.5).  Also records max(bin)+1 per
// channel, the histogram length that hist_gradient_2D uses (lib_image.cc:3674).
__global__ __launch_bounds__(256) void lab_quantize_kernel(const uint32_t* __restrict__ rgb, size_t n,
                                                           const double* __restrict__ gamma_lut,
                                                           uint8_t* __restrict__ qL, uint8_t* __restrict__ qa,
                                                           uint8_t* __restrict__ qb, int* __restrict__ nbins) {
#pragma clang fp contract(off)
  __shared__ int smax[3];
  if (threadIdx.x < 3) smax[threadIdx.x] = 0;
  __syncthreads();
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < n) {
    const uint32_t px = rgb[p];
    const double r = gamma_lut[px & 0xff], g = gamma_lut[(px >> 8) & 0xff], b = gamma_lut[(px >> 16) & 0xff];
    const double X = (0.412453 * r) + (0.357580 * g) + (0.180423 * b);
    const double Y = (0.212671 * r) + (0.715160 * g) + (0.072169 * b);
    const double Z = (0.019334 * r) + (0.119193 * g) + (0.950227 * b);
    const double x = X / 0.950456, y = Y / 1.000000, z = Z / 1.088754;
    const double th = 0.008856;
    const double fx = (x > th) ? pow(x, 1.0 / 3.0) : (7.787 * x + (16.0 / 116.0));
    const double fy = (y > th) ? pow(y, 1.0 / 3.0) : (7.787 * y + (16.0 / 116.0));
    const double fz = (z > th) ? pow(z, 1.0 / 3.0) : (7.787 * z + (16.0 / 116.0));
    const double L = (y > th) ? (116 * pow(y, 1.0 / 3.0) - 16) : (903.3 * y);
    const double A = 500 * (fx - fy), B = 200 * (fy - fz);
    double v[3] = {L / 100.0, (A - (-73)) / (95 - (-73)), (B - (-73)) / (95 - (-73))};
    uint8_t* outs[3] = {qL, qa, qb};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double t = v[c];
      if (t < 0) t = 0;
      else if (t > 1) t = 1;
      int bin = int(floor(t * double(kLabBins)));
      if (bin == kLabBins) bin = kLabBins - 1;
      outs[c][p] = uint8_t(bin);
      atomicMax(&smax[c], bin + 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicMax(&nbins[threadIdx.x], smax[threadIdx.x]);
}

__global__ __launch_bounds__(256) void label_max_kernel(const uint8_t* __restrict__ lab, size_t n, int* __restrict__ nb) {
  __shared__ int smax;
  if (threadIdx.x == 0) smax = 0;
  __syncthreads();
  int m = 0;
  for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += size_t(gridDim.x) * blockDim.x)
    m = max(m, int(lab[p]) + 1);
  atomicMax(&smax, m);
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(nb, smax);
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- half-disc chi-squared ---------------------------------------------------
constexpr int kCueTile = 8;                         // 8x8 pixels per wavefront
constexpr int kCueRmax = 20;
constexpr int kCueTileW = kCueTile + 2 * kCueRmax;  // 48
constexpr int kCueLanes = kCueTile * kCueTile;      // 64
constexpr int kNbMax = 32;                          // largest histogram (textons)
// Histogram rows per slice group: kNbMax bins plus one sink bin that the
// padding entries of the slice-sorted offset lists update and no snapshot reads.
constexpr int kNbStride = kNbMax + 1;
// Sentinel rows below the label tile (value kNbMax = the sink bin): a padding
// entry's offset reaches them from every lane's centre.
constexpr int kCueSentinelRows = kCueTile;
constexpr int kCuePadOffset = (kCueTileW - kCueRmax) * kCueTileW;
// Slice-sorted rings: 5 < d <= 10 and 10 < d <= 20, each list of a slice
// padded to a multiple of kCueBatch entries.
constexpr int kCueRings = 2;
constexpr int kCueBatch = 8;

struct CueParams {
  const uint8_t* plane[4];  // L, a, b, texton labels (H x W, unmirrored)
  const int* nbins;         // [4] histogram lengths
  const int* off;           // disc offsets (LDS tile offset) sorted by radius
  const uint8_t* slice;     // slice index per offset
  int end_by_r[21];         // number of offsets with d^2 <= r^2
  int nslice[4][kNumSlices];  // disc pixels per slice at the snapshot radii 3, 5, 10, 20
  const int* ring_off;      // rings 5 < d <= 10 and 10 < d <= 20, offsets sorted by slice, padded
  int ring_seg[kCueRings * kNumSlices + 1];  // start of (ring, slice) list i in ring_off
  float smooth[3][9];       // half kernels: [0] bg (sigma 2.5, 17 taps), [1] cg (sigma 1.25, 9 taps), [2] none
  float w_mpb[12], w_gpb[12];
  double smooth_d[3][9];    // the same in FP64 (the all-FP64 variant)
  double w_mpb_d[12], w_gpb_d[12];
  void* mpb_all;            // [8][N] (float, or double in the FP64 variant)
  void* gcue;               // [8][N] sum of gPb-weighted cues
  void* cues;               // optional [12][8][N]
  float* feat;              // optional [32][ceil(H/4)][ceil(W/4)]: per channel and orientation, the
                            // scale-weighted mPb contribution averaged over 4x4 cells (FP32 kernel only)
  int W, H;
  int by0;                  // first tile row of this launch (launches may cover row bands)
};

template <typename CT>
__device__ __forceinline__ const CT* smooth_of(const CueParams& P, int k);
template <>
__device__ __forceinline__ const float* smooth_of<float>(const CueParams& P, int k) { return P.smooth[k]; }
template <>
__device__ __forceinline__ const double* smooth_of<double>(const CueParams& P, int k) { return P.smooth_d[k]; }
template <typename CT>
__device__ __forceinline__ CT wmpb_of(const CueParams& P, int c) {
  if constexpr (sizeof(CT) == 8) return P.w_mpb_d[c]; else return P.w_mpb[c];
}
template <typename CT>
__device__ __forceinline__ CT wgpb_of(const CueParams& P, int c) {
  if constexpr (sizeof(CT) == 8) return P.w_gpb_d[c]; else return P.w_gpb[c];
}

// One snapshot: chi-squared of all orientations from the lane's histogram.
// Histogram word [grp][bin][lane] holds the 8-bit counts of slices 4*grp..+3.
// csum[b] (LDS, shared by the wave) = column sums of the truncated smoothing.
//
// The histogram has nb <= NB live bins (nb = largest label + 1 of the
// channel).  Counts of bins nb..NB-1 are zero, so every per-bin quantity built
// from counts (slice masses, U, V and their smoothings) is unchanged by running
// the loops over the compile-time NB; only the chi-squared sum is cut at nb,
// through rT[b] = 0 for b >= nb (smoothing spills mass past nb, and the
// truncated histogram has no such bins).
template <typename CT>
__device__ __forceinline__ CT recip(CT x) {
  if constexpr (sizeof(CT) == 4) return __builtin_amdgcn_rcpf(x);  // 1 ulp
  else return CT(1) / x;
}
// Pass A's mass correction over KH consecutive bins from b0: alpha[s] -=
// (1 - c_b) h_s[b], c_b = csum[b].
template <int KH, typename CT>
__device__ __forceinline__ void chi2_mass_edge(const uint32_t* __restrict__ hist, int lane, int b0,
                                               const CT* __restrict__ csum, CT alpha[kNumSlices]) {
#pragma unroll
  for (int i = 0; i < KH; i++) {
    const int b = b0 + i;
    const CT e = CT(1) - csum[b];
#pragma unroll
    for (int grp = 0; grp < 4; grp++) {
      const uint32_t w = hist[(grp * kNbStride + b) * kCueLanes + lane];
#pragma unroll
      for (int q = 0; q < 4; q++) alpha[4 * grp + q] = fma_t(-e, CT((w >> (8 * q)) & 0xffu), alpha[4 * grp + q]);
    }
  }
}

// ns: the disc pixels of each slice at this radius (every label of a channel
// is below its nb, so the counts of a slice sum to ns[s]).
template <int NB, int KH, typename CT>
__device__ __forceinline__ void chi2_snapshot(const uint32_t* __restrict__ hist, int lane, int nb,
                                           const CT* __restrict__ g, const CT* __restrict__ csum,
                                           const int* __restrict__ ns, CT out[kNumOri]) {
  // pass A: per-slice mass after smoothing, a_s = 1 / mass.  The smoothing
  // kernel sums to 1, so a bin whose taps all fall inside [0, nb) keeps its
  // whole count: mass_s = ns[s] - sum_b (1 - c_b) h_s[b] over the bins within
  // KH of either end only (all bins when the two ends overlap; none without
  // smoothing, where the mass is the pixel count itself)
  CT alpha[kNumSlices];
#pragma unroll
  for (int s = 0; s < kNumSlices; s++) alpha[s] = CT(ns[s]);
  if (KH > 0) {
    if (nb >= 2 * KH) {  // uniform
      chi2_mass_edge<KH, CT>(hist, lane, 0, csum, alpha);
      chi2_mass_edge<KH, CT>(hist, lane, nb - KH, csum, alpha);
    } else {
#pragma unroll 1
      for (int b = 0; b < nb; b++) chi2_mass_edge<1, CT>(hist, lane, b, csum, alpha);
    }
  }
#pragma unroll
  for (int s = 0; s < kNumSlices; s++) alpha[s] = alpha[s] > CT(0) ? recip(alpha[s]) : CT(0);
  // The passes below read the histogram words again rather than keep pass
  // A's 4 * NB words (or their 16 * NB converted counts) live in registers.
  asm volatile("" ::: "memory");
  // pass B: U_T = sum_s a_s h_s, V_0 = right - left for orientation 0
  CT T[NB], V[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    CT l = CT(0), r = CT(0);  // the halves of orientation 0: slices 0..7 and 8..15
#pragma unroll
    for (int grp = 0; grp < 4; grp++) {
      const uint32_t w = hist[(grp * kNbStride + b) * kCueLanes + lane];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int s = 4 * grp + q;
        const CT c = CT((w >> (8 * q)) & 0xffu);
        if (s < kNumOri)
          l = fma_t(alpha[s], c, l);
        else
          r = fma_t(alpha[s], c, r);
      }
    }
    T[b] = l + r;  // U_T for now, smoothed below
    V[b] = r - l;
  }
  if (KH > 0) {
    CT U[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) U[b] = T[b];
#pragma unroll
    for (int b = 0; b < NB; b++) {
      CT acc = CT(0);
#pragma unroll
      for (int k = -KH; k <= KH; k++) {
        const int n = b + k;
        if (n >= 0 && n < NB) acc = fma_t(g[k < 0 ? -k : k], U[n], acc);
      }
      T[b] = acc;
    }
  }
  CT rT[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) rT[b] = (b < nb && T[b] > CT(0)) ? recip(T[b]) : CT(0);
  asm volatile("" ::: "memory");
  // orientations in two rolled rounds of four unrolled ones: the byte of a
  // slice within its histogram word is then a constant
#pragma unroll 1
  for (int oh = 0; oh < kNumOri; oh += 4)
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int o = oh + q;
    CT acc = CT(0);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      CT d = V[b];
      if (KH > 0) {
        d = CT(0);
#pragma unroll
        for (int k = -KH; k <= KH; k++) {
          const int n = b + k;
          if (n >= 0 && n < NB) d = fma_t(g[k < 0 ? -k : k], V[n], d);
        }
      }
      acc = fma_t(d * d, rT[b], acc);
    }
    out[o] = CT(0.5) * acc;
    if (o + 1 < kNumOri) {
      // orientation o + 1: slice o moves from the right half to the left,
      // slice o + 8 from the left to the right
      const CT a0 = CT(2) * (oh ? alpha[4 + q] : alpha[q]);
      const CT a8 = -CT(2) * (oh ? alpha[kNumOri + 4 + q] : alpha[kNumOri + q]);
      const uint32_t* h0 = hist + ((oh >> 2) * kNbStride) * kCueLanes + lane;
      const uint32_t* h8 = hist + (((oh >> 2) + 2) * kNbStride) * kCueLanes + lane;
#pragma unroll
      for (int b = 0; b < NB; b++) {
        const CT c0 = CT((h0[b * kCueLanes] >> (8 * q)) & 0xffu);
        const CT c8 = CT((h8[b * kCueLanes] >> (8 * q)) & 0xffu);
        V[b] = fma_t(a8, c8, fma_t(a0, c0, V[b]));
      }
    }
  }
}

template <int NB, int KH, typename CT>
__device__ void cue_channel(const CueParams& P, int ch, int ri0, const int* radii, uint8_t* tile, uint32_t* hist,
                            CT* csum, int lane, int px, int py, bool valid, CT mpb[kNumOri], CT gc[kNumOri]) {
  const int W = P.W, H = P.H;
  const int gx0 = int(blockIdx.x) * kCueTile - kCueRmax, gy0 = (int(blockIdx.y) + P.by0) * kCueTile - kCueRmax;
  const uint8_t* plane = P.plane[ch];
  for (int i = lane; i < kCueTileW * kCueTileW; i += kCueLanes) {
    const int ty = i / kCueTileW, tx = i - ty * kCueTileW;
    int gx = gx0 + tx, gy = gy0 + ty;
    // half-sample symmetric border (border_mirror_2D), periodic beyond the image
    if (W == 1) gx = 0; else { const int p2 = 2 * W; gx %= p2; if (gx < 0) gx += p2; if (gx >= W) gx = p2 - 1 - gx; }
    if (H == 1) gy = 0; else { const int p2 = 2 * H; gy %= p2; if (gy < 0) gy += p2; if (gy >= H) gy = p2 - 1 - gy; }
    tile[i] = plane[size_t(gy) * W + gx];
  }
  for (int i = 0; i < 4 * kNbStride; i++) hist[i * kCueLanes + lane] = 0u;
  const int nb = P.nbins[ch];
  const CT* g = smooth_of<CT>(P, KH == 8 ? 0 : (KH == 4 ? 1 : 2));
  if (lane < kNbMax) {
    CT s = 0;
    for (int k = -KH; k <= KH; k++) {
      const int n = lane - k;
      if (n >= 0 && n < nb) s += g[k < 0 ? -k : k];
    }
    csum[lane] = (lane < nb) ? s : CT(0);
  }
  __syncthreads();
  const int ty = lane / kCueTile, tx = lane - ty * kCueTile;
  const uint8_t* centre = tile + (ty + kCueRmax) * kCueTileW + tx + kCueRmax;
  const __attribute__((address_space(4))) int* offc = (const __attribute__((address_space(4))) int*)P.off;
  const __attribute__((address_space(4))) uint8_t* slc = (const __attribute__((address_space(4))) uint8_t*)P.slice;
  CT chf[kNumOri];
#pragma unroll
  for (int o = 0; o < kNumOri; o++) chf[o] = CT(0);
  int i = 0;
  for (int ri = 0; ri < 3; ri++) {
    if (radii[ri] > 5) {
      // outer rings from the slice-sorted lists: within a slice the histogram
      // group is a constant offset and the increment a constant
      const int ring = radii[ri] == 10 ? 0 : 1;
      const __attribute__((address_space(4))) int* roff = (const __attribute__((address_space(4))) int*)P.ring_off;
#pragma unroll
      for (int sl = 0; sl < kNumSlices; sl++) {
        uint32_t* hs = hist + (sl >> 2) * kNbStride * kCueLanes + lane;
        const uint32_t inc = 1u << (8 * (sl & 3));
        const int e = P.ring_seg[ring * kNumSlices + sl + 1];
        for (int j = P.ring_seg[ring * kNumSlices + sl]; j < e; j += kCueBatch) {
          uint32_t lab[kCueBatch];
#pragma unroll
          for (int k = 0; k < kCueBatch; k++) lab[k] = centre[roff[j + k]];
#pragma unroll
          for (int k = 0; k < kCueBatch; k++) atomicAdd(&hs[lab[k] * kCueLanes], inc);
        }
      }
    } else {
    const int end = P.end_by_r[radii[ri]];
    // 8 label reads are issued before their 8 histogram updates (read and
    // update one at a time, each read waited behind the previous update: the
    // compiler cannot separate the two LDS arrays here); the uniform offset
    // and slice tables come through the scalar cache, the next batch's
    // entries loaded before this batch's LDS work
    if (i + 8 <= end) {
      int off_n[8], sl_n[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        off_n[k] = offc[i + k];
        sl_n[k] = slc[i + k];
      }
      for (; i + 8 <= end; i += 8) {
        int off_c[8], sl_c[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          off_c[k] = off_n[k];
          sl_c[k] = sl_n[k];
        }
        if (i + 16 <= end) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            off_n[k] = offc[i + 8 + k];
            sl_n[k] = slc[i + 8 + k];
          }
        }
        uint32_t lab[8];
#pragma unroll
        for (int k = 0; k < 8; k++) lab[k] = centre[off_c[k]];
#pragma unroll
        for (int k = 0; k < 8; k++)
          atomicAdd(&hist[((sl_c[k] >> 2) * kNbStride + lab[k]) * kCueLanes + lane], 1u << (8 * (sl_c[k] & 3)));
      }
    }
    for (; i < end; i++) {
      const int s = P.slice[i];
      const uint32_t lab = centre[P.off[i]];
      atomicAdd(&hist[((s >> 2) * kNbStride + lab) * kCueLanes + lane], 1u << (8 * (s & 3)));
    }
    }
    CT chi[kNumOri];
    const int rsel = radii[ri] == 3 ? 0 : (radii[ri] == 5 ? 1 : (radii[ri] == 10 ? 2 : 3));
    chi2_snapshot<NB, KH, CT>(hist, lane, nb, g, csum, P.nslice[rsel], chi);
    const int cue = ri0 + ri;
    if (valid) {
      const size_t N = size_t(W) * H, p = size_t(py) * W + px;
#pragma unroll
      for (int o = 0; o < kNumOri; o++) {
        mpb[o] = fma_t(wmpb_of<CT>(P, cue), chi[o], mpb[o]);
        gc[o] = fma_t(wgpb_of<CT>(P, cue), chi[o], gc[o]);
        chf[o] = fma_t(wmpb_of<CT>(P, cue), chi[o], chf[o]);
        if (P.cues) static_cast<CT*>(P.cues)[(size_t(cue) * kNumOri + o) * N + p] = chi[o];
      }
    }
  }
  if constexpr (sizeof(CT) == 4) {
    if (P.feat) {
      // 4x4 cell mean: the cell's 16 lanes differ in lane bits 0, 1 (x) and 3, 4 (y)
      const int Wq = (W + 3) / 4, Hq = (H + 3) / 4;
      float cnt = valid ? 1.f : 0.f;
      for (int m = 1; m <= 16; m <<= 1) {
        if (m == 4) continue;
        cnt += __shfl_xor(cnt, m, 64);
      }
#pragma unroll
      for (int o = 0; o < kNumOri; o++) {
        float v = valid ? float(chf[o]) : 0.f;
        for (int m = 1; m <= 16; m <<= 1) {
          if (m == 4) continue;
          v += __shfl_xor(v, m, 64);
        }
        if ((lane & 0x1b) == 0 && cnt > 0.f)
          P.feat[(size_t(ch) * kNumOri + o) * size_t(Hq) * Wq + size_t(py >> 2) * Wq + (px >> 2)] = v / cnt;
      }
    }
  }
  __syncthreads();
}

template <typename CT>
__global__ __launch_bounds__(64) void cues_kernel(CueParams P) {
  __shared__ uint8_t tile[kCueTileW * (kCueTileW + kCueSentinelRows)];
  __shared__ uint32_t hist[4 * kNbStride * kCueLanes];
  __shared__ CT csum[kNbMax];
  const int lane = threadIdx.x;
  for (int i = lane; i < kCueSentinelRows * kCueTileW; i += kCueLanes) tile[kCueTileW * kCueTileW + i] = uint8_t(kNbMax);
  const int px = blockIdx.x * kCueTile + (lane % kCueTile), py = (blockIdx.y + P.by0) * kCueTile + (lane / kCueTile);
  const bool valid = px < P.W && py < P.H;
  CT mpb[kNumOri], gc[kNumOri];
#pragma unroll
  for (int o = 0; o < kNumOri; o++) mpb[o] = gc[o] = CT(0);
  static constexpr int rbg[3] = {3, 5, 10}, rcg[3] = {5, 10, 20};
  cue_channel<kLabBins, 8, CT>(P, 0, 0, rbg, tile, hist, csum, lane, px, py, valid, mpb, gc);
  cue_channel<kLabBins, 4, CT>(P, 1, 3, rcg, tile, hist, csum, lane, px, py, valid, mpb, gc);
  cue_channel<kLabBins, 4, CT>(P, 2, 6, rcg, tile, hist, csum, lane, px, py, valid, mpb, gc);
  cue_channel<kTextonK, 0, CT>(P, 3, 9, rcg, tile, hist, csum, lane, px, py, valid, mpb, gc);
  if (valid) {
    const size_t N = size_t(P.W) * P.H, p = size_t(py) * P.W + px;
#pragma unroll
    for (int o = 0; o < kNumOri; o++) {
      static_cast<CT*>(P.mpb_all)[o * N + p] = mpb[o];
      static_cast<CT*>(P.gcue)[o * N + p] = gc[o];
    }
  }
}

// ---- oriented nonmax ---------------------------------------------------------
template <typename VT>
__global__ __launch_bounds__(256) void max_ori_kernel(const VT* __restrict__ mpb_all, size_t n,
                                                      VT* __restrict__ a, uint8_t* __restrict__ k) {
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  VT mv = -VT(__builtin_huge_val());
  int mi = 0;
#pragma unroll
  for (int o = 0; o < kNumOri; o++) {
    const VT v = mpb_all[o * n + p];
    if (v > mv) {
      mv = v;
      mi = o;
    }
  }
  a[p] = mv < VT(0) ? VT(0) : mv;
  k[p] = uint8_t(mi);
}

struct NmsTable {
  int branch[kNumOri];
  double d[kNumOri];
  double cosd[kNumOri][kNumOri];
};

// nonmax_oriented_2D in FP64 on the VT inputs, then the multiscalePb clip.
template <typename VT>
__global__ __launch_bounds__(256) void nonmax_kernel(const VT* __restrict__ a, const uint8_t* __restrict__ k, int W,
                                                     int H, NmsTable T, VT* __restrict__ nmax) {
#pragma clang fp contract(off)
  const int x = blockIdx.x * blockDim.x + threadIdx.x;  // column (lib_image y)
  const int y = blockIdx.y;                             // row (lib_image x)
  if (x >= W) return;
  const size_t n = size_t(y) * W + x;
  const int kc = k[n];
  const int br = T.branch[kc];
  const double d = T.d[kc];
  bool ok0 = false, ok1 = false;
  long i0a = 0, i0b = 0, i1a = 0, i1b = 0;
  const long sy = W;
  switch (br) {
    case 0:
      ok0 = y > 0; ok1 = y < H - 1;
      i0a = i0b = long(n) - sy; i1a = i1b = long(n) + sy;
      break;
    case 1:
      ok0 = y > 0 && x > 0; ok1 = y < H - 1 && x < W - 1;
      i0a = long(n) - sy; i0b = i0a - 1; i1a = long(n) + sy; i1b = i1a + 1;
      break;
    case 2:
      ok0 = y > 0 && x > 0; ok1 = y < H - 1 && x < W - 1;
      i0a = long(n) - 1; i0b = i0a - sy; i1a = long(n) + 1; i1b = i1a + sy;
      break;
    case 3:
      ok0 = x > 0; ok1 = x < W - 1;
      i0a = i0b = long(n) - 1; i1a = i1b = long(n) + 1;
      break;
    case 4:
      ok0 = y < H - 1 && x > 0; ok1 = y > 0 && x < W - 1;
      i0a = long(n) - 1; i0b = i0a + sy; i1a = long(n) + 1; i1b = i1a - sy;
      break;
    default:
      ok0 = y < H - 1 && x > 0; ok1 = y > 0 && x < W - 1;
      i0a = long(n) + sy; i0b = i0a - 1; i1a = long(n) - sy; i1b = i1a + 1;
      break;
  }
  double res = 0;
  if (ok0 && ok1) {
    const double v = a[n];
    const double v0 = (1.0 - d) * double(a[i0a]) * T.cosd[kc][k[i0a]] + d * double(a[i0b]) * T.cosd[kc][k[i0b]];
    const double v1 = (1.0 - d) * double(a[i1a]) * T.cosd[kc][k[i1a]] + d * double(a[i1b]) * T.cosd[kc][k[i1b]];
    if (v > v0 && v > v1) res = v;
  }
  if (1.2 * res > 1) res = 1;
  if (res < 0) res = 0;
  nmax[n] = VT(res);
}

// ---- host side -------------------------------------------------------------------
namespace {
struct CueTables {
  DevBuf<int> off;
  DevBuf<int> ring_off;
  int ring_seg[kCueRings * kNumSlices + 1];
  DevBuf<uint8_t> slice;
  DevBuf<double> gamma;
  int end_by_r[21];
  int nslice[4][kNumSlices];
  float smooth[3][9];
  double smooth_d[3][9];
  NmsTable nms;
};
std::map<int, std::unique_ptr<CueTables>> g_cue_tables;  // per device
std::mutex g_cue_mu;
}  // namespace

static CueTables& cue_tables(Context& c) {
  std::lock_guard<std::mutex> lk(g_cue_mu);
  auto& slot = g_cue_tables[c.device];
  if (slot) return *slot;
  auto t = std::make_unique<CueTables>();
  std::vector<std::array<int, 3>> offs;  // d2, dy, dx
  for (int dy = -kCueRmax; dy <= kCueRmax; dy++)
    for (int dx = -kCueRmax; dx <= kCueRmax; dx++)
      if (dx * dx + dy * dy <= kCueRmax * kCueRmax) offs.push_back({dx * dx + dy * dy, dy, dx});
  std::stable_sort(offs.begin(), offs.end(), [](const auto& a, const auto& b) { return a[0] < b[0]; });
  std::vector<int> off(offs.size());
  std::vector<uint8_t> sl(offs.size());
  for (size_t i = 0; i < offs.size(); i++) {
    off[i] = offs[i][1] * kCueTileW + offs[i][2];
    sl[i] = uint8_t(slice_of_offset(offs[i][1], offs[i][2]));
  }
  for (int r = 0; r <= kCueRmax; r++) {
    int e = 0;
    while (e < int(offs.size()) && offs[size_t(e)][0] <= r * r) e++;
    t->end_by_r[r] = e;
  }
  {
    const int rs[4] = {3, 5, 10, 20};
    for (int k = 0; k < 4; k++) {
      for (int q = 0; q < kNumSlices; q++) t->nslice[k][q] = 0;
      for (size_t i = 0; i < offs.size(); i++)
        if (offs[i][0] <= rs[k] * rs[k]) t->nslice[k][sl[i]]++;
    }
  }
  // the two outer rings, by slice, each slice's list padded to kCueBatch
  // entries with the sentinel offset
  {
    std::vector<int> roff;
    int n = 0;
    for (int ring = 0; ring < kCueRings; ring++) {
      const int r0 = ring == 0 ? 5 : 10, r1 = ring == 0 ? 10 : 20;
      for (int sidx = 0; sidx < kNumSlices; sidx++) {
        t->ring_seg[n++] = int(roff.size());
        for (size_t i = 0; i < offs.size(); i++)
          if (offs[i][0] > r0 * r0 && offs[i][0] <= r1 * r1 && sl[i] == sidx) roff.push_back(off[i]);
        while (roff.size() % kCueBatch) roff.push_back(kCuePadOffset);
      }
    }
    t->ring_seg[n] = int(roff.size());
    t->ring_off.alloc(roff.size());
    GPB_HIP(hipMemcpy(t->ring_off.p, roff.data(), roff.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  t->off.alloc(off.size());
  t->slice.alloc(sl.size());
  GPB_HIP(hipMemcpy(t->off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
  GPB_HIP(hipMemcpy(t->slice.p, sl.data(), sl.size(), hipMemcpyHostToDevice));
  std::vector<double> gam(256);
  for (int i = 0; i < 256; i++) gam[size_t(i)] = std::pow(double(i) / 255.0, 2.5);
  t->gamma.alloc(256);
  GPB_HIP(hipMemcpy(t->gamma.p, gam.data(), 256 * sizeof(double), hipMemcpyHostToDevice));
  const std::vector<double> gbg = gaussian_1d(0.1 * kLabBins), gcg = gaussian_1d(0.05 * kLabBins);
  if (gbg.size() != 17 || gcg.size() != 9) throw std::logic_error("unexpected smoothing kernel size");
  for (int k = 0; k < 9; k++) {
    t->smooth[0][k] = float(gbg[size_t(8 + k)]);
    t->smooth[1][k] = k < 5 ? float(gcg[size_t(4 + k)]) : 0.f;
    t->smooth[2][k] = k == 0 ? 1.f : 0.f;
    t->smooth_d[0][k] = gbg[size_t(8 + k)];
    t->smooth_d[1][k] = k < 5 ? gcg[size_t(4 + k)] : 0.0;
    t->smooth_d[2][k] = k == 0 ? 1.0 : 0.0;
  }
  const NonmaxTable& nt = nonmax_table();
  for (int k = 0; k < kNumOri; k++) {
    t->nms.branch[k] = nt.branch[k];
    t->nms.d[k] = nt.d[k];
    for (int j = 0; j < kNumOri; j++) t->nms.cosd[k][j] = nt.cosdiff[k][j];
  }
  slot = std::move(t);
  return *slot;
}

template <typename CT>
static void local_cues_t(Context& c, const uint32_t* d_rgb, const uint8_t* d_text, int W, int H, CT* d_mpb_all,
                         CT* d_gcue, CT* d_cues, float* d_feat) {
  CueTables& t = cue_tables(c);
  const size_t N = size_t(W) * H;
  uint8_t* q = c.ws<uint8_t>("cue.q", 3 * N);
  int* nb = c.ws<int>("cue.nb", 4);
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, c.stream, nb, 4, 0);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(lab_quantize_kernel, dim3(div_up(N, 256)), dim3(256), 0, c.stream, d_rgb, N, t.gamma.p, q, q + N,
                     q + 2 * N, nb);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(label_max_kernel, dim3(std::min<unsigned>(1024, div_up(N, 256))), dim3(256), 0, c.stream, d_text, N,
                     nb + 3);
  GPB_LAUNCH_CHECK();
  CueParams P{};
  P.plane[0] = q;
  P.plane[1] = q + N;
  P.plane[2] = q + 2 * N;
  P.plane[3] = d_text;
  P.nbins = nb;
  P.off = t.off.p;
  P.slice = t.slice.p;
  P.ring_off = t.ring_off.p;
  for (int q = 0; q <= kCueRings * kNumSlices; q++) P.ring_seg[q] = t.ring_seg[q];
  for (int r = 0; r <= kCueRmax; r++) P.end_by_r[r] = t.end_by_r[r];
  for (int k = 0; k < 4; k++)
    for (int q = 0; q < kNumSlices; q++) P.nslice[k][q] = t.nslice[k][q];
  for (int a = 0; a < 3; a++)
    for (int k = 0; k < 9; k++) {
      P.smooth[a][k] = t.smooth[a][k];
      P.smooth_d[a][k] = t.smooth_d[a][k];
    }
  for (int i = 0; i < kNumCues; i++) {
    P.w_mpb[i] = float(kMPbWeights[i]);
    P.w_gpb[i] = float(kGPbWeights[i]);
    P.w_mpb_d[i] = kMPbWeights[i];
    P.w_gpb_d[i] = kGPbWeights[i];
  }
  P.mpb_all = d_mpb_all;
  P.gcue = d_gcue;
  P.cues = d_cues;
  P.feat = d_feat;
  P.W = W;
  P.H = H;
  KScope ks(c, kKCues, double(N) * (4 + 16 * sizeof(CT) + (d_cues ? 96 * sizeof(CT) : 0)) + (d_feat ? 8.0 * N : 0.0));
  const int ty = int(div_up(H, kCueTile));
  // verbose runs launch row bands of 64 tile rows, each followed by a progress line
  const int band = c.verbose ? 64 : ty;
  for (int y0 = 0; y0 < ty; y0 += band) {
    P.by0 = y0;
    hipLaunchKernelGGL(cues_kernel<CT>, dim3(div_up(W, kCueTile), std::min(band, ty - y0)), dim3(kCueLanes), 0,
                       c.stream, P);
    GPB_LAUNCH_CHECK();
    if (c.verbose && ty > band) {
      GPB_HIP(hipStreamSynchronize(c.stream));
      std::fprintf(stderr, "gpb: local cues (%s) %dx%d: rows %d of %d\n", sizeof(CT) == 8 ? "FP64" : "FP32", W, H,
                   std::min(H, (y0 + band) * kCueTile), H);
    }
  }
}

void launch_local_cues(Context& c, const uint32_t* d_rgb, const uint8_t* d_text, int W, int H, float* d_mpb_all,
                       float* d_gcue, float* d_cues, float* d_feat) {
  local_cues_t<float>(c, d_rgb, d_text, W, H, d_mpb_all, d_gcue, d_cues, d_feat);
}

void launch_local_cues_f64(Context& c, const uint32_t* d_rgb, const uint8_t* d_text, int W, int H, double* d_mpb_all,
                           double* d_gcue, double* d_cues) {
  local_cues_t<double>(c, d_rgb, d_text, W, H, d_mpb_all, d_gcue, d_cues, nullptr);
}

void launch_lab_quantize(Context& c, const uint32_t* d_rgb, size_t N, uint8_t* d_q /* 3N */, int* d_nb /* 3 */) {
  CueTables& t = cue_tables(c);
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, c.stream, d_nb, 3, 0);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(lab_quantize_kernel, dim3(div_up(N, 256)), dim3(256), 0, c.stream, d_rgb, N, t.gamma.p, d_q,
                     d_q + N, d_q + 2 * N, d_nb);
  GPB_LAUNCH_CHECK();
}

template <typename VT>
static void nonmax_t(Context& c, const VT* d_mpb_all, int W, int H, VT* d_nmax) {
  CueTables& t = cue_tables(c);
  const size_t N = size_t(W) * H;
  VT* a = c.ws<VT>(sizeof(VT) == 8 ? "nms.a64" : "nms.a", N);
  uint8_t* k = c.ws<uint8_t>("nms.k", N);
  KScope ks(c, kKNonmax, double(N) * (10 * sizeof(VT) + 6));
  hipLaunchKernelGGL(max_ori_kernel<VT>, dim3(div_up(N, 256)), dim3(256), 0, c.stream, d_mpb_all, N, a, k);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(nonmax_kernel<VT>, dim3(div_up(W, 256), H), dim3(256), 0, c.stream, a, k, W, H, t.nms, d_nmax);
  GPB_LAUNCH_CHECK();
}

void launch_nonmax(Context& c, const float* d_mpb_all, int W, int H, float* d_nmax) {
  nonmax_t<float>(c, d_mpb_all, W, H, d_nmax);
}

void launch_nonmax_f64(Context& c, const double* d_mpb_all, int W, int H, double* d_nmax) {
  nonmax_t<double>(c, d_mpb_all, W, H, d_nmax);
}

void launch_nonmax_scratch(hipStream_t stream, const float* d_mpb_all, int W, int H, float* a, uint8_t* k,
                           float* d_nmax) {
  NmsTable T;
  const NonmaxTable& nt = nonmax_table();
  for (int o = 0; o < kNumOri; o++) {
    T.branch[o] = nt.branch[o];
    T.d[o] = nt.d[o];
    for (int j = 0; j < kNumOri; j++) T.cosd[o][j] = nt.cosdiff[o][j];
  }
  const size_t N = size_t(W) * H;
  hipLaunchKernelGGL(max_ori_kernel<float>, dim3(div_up(N, 256)), dim3(256), 0, stream, d_mpb_all, N, a, k);
  GPB_LAUNCH_CHECK();
  hipLaunchKernelGGL(nonmax_kernel<float>, dim3(div_up(W, 256), H), dim3(256), 0, stream, a, k, W, H, T, d_nmax);
  GPB_LAUNCH_CHECK();
}

}  // namespace gpb

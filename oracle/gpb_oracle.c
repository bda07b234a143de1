/*
 * gpb_oracle.c — CPU restatement of the reference gPb CPU path, used ONLY as a
 * test oracle and as the CPU baseline of the benchmark.  Nothing in the
 * product (libgpb.so) links, loads or calls this code.
 *
 * Each function restates the reference routine named in its comment
 * (paths relative to the yanyp/UCM-MPI-GPU tree).  Arithmetic is FP64 except
 * where the reference path is FP32 (Schmid responses and texton distances,
 * whose accumulation order is declared below so that labels are
 * bit-reproducible).  Deviations from the reference, each deliberate:
 *   - colour channels are read as R, G, B (the CPU path hands OpenCV's B, G, R
 *     planes to routines written for R, G, B: det_mPb.cpp:11 ->
 *     mex_pb_parts_final_selected.cpp:104-119; the GPU path decodes R, G, B);
 *   - texton labels are the true argmin (SPFeatures.cpp:328-339 never updates
 *     minDistance, so every label is 31);
 *   - borders wider than the image reflect periodically (the reference reads
 *     out of bounds there);
 *   - nonmax_channels starts each pixel at -inf with its index reset
 *     (nonmax_channels.cpp:17-34 carries both across pixels);
 *   - the operator is the full symmetric D - W (spectralPb.cpp:199-209 keeps
 *     only the lower triangle and multiplies with a general product);
 *   - eigenpairs are the K smallest sorted ascending and the trivial one is
 *     dropped (SymEigsSolver sorts largest first, SymEigsSolver.h:563-580);
 *   - sPb filters follow the GPU construction (gaussian_2D,
 *     damascene/src/textons/filters.cpp:264-341), since oeFilter.cpp:21-196
 *     reads uninitialised memory.
 * Parity status of each function is recorded in the test files that pin it.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NORI 8
#define NF 13
#define SUP 49
#define HSUP 24
#define KCB 32
static const double REF_PI = 3.1415926535; /* include/UCM_CONSTANTS.h:4 */

static int refl(int i, int n) { /* half-sample symmetric, periodic */
  if (n == 1) return 0;
  int p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - 1 - i;
}
static int refl101(int i, int n) { /* whole-sample symmetric, periodic */
  if (n == 1) return 0;
  int p = 2 * n - 2;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - i;
}

/* ---- a2: grey for the Schmid path: OpenCV BGR2GRAY on 8-bit data
 * (SPFeatures.cpp:27): Y = (4899 R + 9617 G + 1868 B + 2^13) >> 14. */
int orc_grey_u8(uint32_t px) {
  int r = px & 0xff, g = (px >> 8) & 0xff, b = (px >> 16) & 0xff;
  return (r * 4899 + g * 9617 + b * 1868 + 8192) >> 14;
}

/* ---- a6: Schmid bank, SPFeatures::makefilter (SPFeatures.cpp:117-165),
 * makeSfilters order (:167-183).  out[f][i][j], i = row (y), j = col (x). */
void orc_schmid_bank(double* out) {
  static const int st[NF][2] = {{2, 1}, {4, 1}, {4, 2}, {6, 1}, {6, 2}, {6, 3}, {8, 1},
                                {8, 2}, {8, 3}, {10, 1}, {10, 2}, {10, 3}, {10, 4}};
  for (int f = 0; f < NF; f++) {
    int sigma = st[f][0], tau = st[f][1];
    double* F = out + (size_t)f * SUP * SUP;
    double fsum = 0, fabs_sum = 0;
    for (int i = 0; i < SUP; i++)
      for (int j = 0; j < SUP; j++) {
        double x = j - HSUP, y = i - HSUP;
        double r = sqrt(x * x + y * y);
        double v = cos(r * (REF_PI * tau / sigma));
        v *= exp(-r * r / (2 * sigma * sigma));
        F[i * SUP + j] = v;
        fsum += v;
        fabs_sum += fabs(v);
      }
    double mean = fsum / (SUP * SUP);
    for (int n = 0; n < SUP * SUP; n++) F[n] = (F[n] - mean) / fabs_sum;
  }
}

/* Direct FP64 correlation of the grey image with the 13 kernels at one pixel
 * (copyMakeBorder REFLECT 24 + filter2D, SPFeatures.cpp:285,299-305). */
void orc_schmid_direct(const uint32_t* packed, int W, int H, int x, int y, double* r13) {
  static double bank[NF * SUP * SUP];
  static int ready = 0;
  if (!ready) {
    orc_schmid_bank(bank);
    ready = 1;
  }
  for (int f = 0; f < NF; f++) {
    double s = 0;
    for (int i = 0; i < SUP; i++)
      for (int j = 0; j < SUP; j++) {
        int gy = refl(y + i - HSUP, H), gx = refl(x + j - HSUP, W);
        s += bank[(size_t)f * SUP * SUP + i * SUP + j] * (double)orc_grey_u8(packed[(size_t)gy * W + gx]);
      }
    r13[f] = s;
  }
}

/* ---- a7/a8: texton labels with the declared FP32 order.
 * Radial classes c sorted by dx^2+dy^2; S_c = exact sum of grey values of the
 * class members; r_f = fmaf(w_f[c], S_c, r_f) over c ascending, from 0.
 * Distances d_k = sum_f fmaf((r_f - C[f][k])^2) over f ascending; label =
 * first k with the smallest d_k. */
void orc_texton(const uint32_t* packed, int W, int H, const float* codebook, uint8_t* labels, float* resp) {
  static double bank[NF * SUP * SUP];
  static int cls_of[SUP * SUP]; /* class index per tap */
  static float wcls[NF][SUP * SUP];
  static int ncls = 0;
  if (!ncls) {
    orc_schmid_bank(bank);
    int used[2 * HSUP * HSUP + 1];
    for (int q = 0; q <= 2 * HSUP * HSUP; q++) used[q] = -1;
    for (int q = 0; q <= 2 * HSUP * HSUP; q++) /* ascending radius */
      for (int i = 0; i < SUP; i++)
        for (int j = 0; j < SUP; j++) {
          int dy = i - HSUP, dx = j - HSUP;
          if (dx * dx + dy * dy != q) continue;
          if (used[q] < 0) {
            used[q] = ncls++;
            for (int f = 0; f < NF; f++) wcls[f][used[q]] = (float)bank[(size_t)f * SUP * SUP + i * SUP + j];
          }
          cls_of[i * SUP + j] = used[q];
        }
  }
  float* S = (float*)malloc(sizeof(float) * ncls);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      for (int c = 0; c < ncls; c++) S[c] = 0.f;
      for (int i = 0; i < SUP; i++)
        for (int j = 0; j < SUP; j++) {
          int gy = refl(y + i - HSUP, H), gx = refl(x + j - HSUP, W);
          S[cls_of[i * SUP + j]] += (float)orc_grey_u8(packed[(size_t)gy * W + gx]);
        }
      float r[NF];
      for (int f = 0; f < NF; f++) {
        float a = 0.f;
        for (int c = 0; c < ncls; c++) a = fmaf(wcls[f][c], S[c], a);
        r[f] = a;
        if (resp) resp[(size_t)f * W * H + (size_t)y * W + x] = a;
      }
      float best = INFINITY;
      int bk = 0;
      for (int k = 0; k < KCB; k++) {
        float acc = 0.f;
        for (int f = 0; f < NF; f++) {
          float d = r[f] - codebook[f * KCB + k];
          acc = fmaf(d, d, acc);
        }
        if (acc < best) {
          best = acc;
          bk = k;
        }
      }
      labels[(size_t)y * W + x] = (uint8_t)bk;
    }
  free(S);
}

/* ---- a3/a4: gamma 2.5, RGB->XYZ->Lab (D65), normalise, quantise to 25 bins
 * (lib_image.cc:123-138, 172-185, 216-245, 143-167, 2315-2344).  Input values
 * are channel/255 (globalPb.cpp:20).  Output bins[3][N]. */
void orc_lab_quantize(const uint32_t* packed, size_t n, int nbins, uint8_t* bins, double* lab) {
  for (size_t p = 0; p < n; p++) {
    double r = (packed[p] & 0xff) / 255.0, g = ((packed[p] >> 8) & 0xff) / 255.0, b = ((packed[p] >> 16) & 0xff) / 255.0;
    r = pow(r, 2.5);
    g = pow(g, 2.5);
    b = pow(b, 2.5);
    double X = (0.412453 * r) + (0.357580 * g) + (0.180423 * b);
    double Y = (0.212671 * r) + (0.715160 * g) + (0.072169 * b);
    double Z = (0.019334 * r) + (0.119193 * g) + (0.950227 * b);
    double x = X / 0.950456, y = Y / 1.000000, z = Z / 1.088754;
    const double th = 0.008856;
    double fx = (x > th) ? pow(x, 1.0 / 3.0) : (7.787 * x + (16.0 / 116.0));
    double fy = (y > th) ? pow(y, 1.0 / 3.0) : (7.787 * y + (16.0 / 116.0));
    double fz = (z > th) ? pow(z, 1.0 / 3.0) : (7.787 * z + (16.0 / 116.0));
    double L = (y > th) ? (116 * pow(y, 1.0 / 3.0) - 16) : (903.3 * y);
    double A = 500 * (fx - fy), B = 200 * (fy - fz);
    double v[3];
    v[0] = L / 100.0;
    v[1] = (A - (-73)) / (95 - (-73));
    v[2] = (B - (-73)) / (95 - (-73));
    for (int c = 0; c < 3; c++) {
      if (v[c] < 0) v[c] = 0;
      else if (v[c] > 1) v[c] = 1;
      unsigned long bin = (unsigned long)floor(v[c] * (double)nbins);
      if (bin == (unsigned long)nbins) bin = nbins - 1;
      bins[(size_t)c * n + p] = (uint8_t)bin;
      if (lab) lab[(size_t)c * n + p] = v[c];
    }
  }
}

/* ---- a5: half-sample symmetric border (lib_image border_mirror_2D,
 * lib_image.cc:308-370).  Generic element size. */
void orc_mirror_u8(const uint8_t* src, int W, int H, int b, uint8_t* dst) {
  int Wm = W + 2 * b, Hm = H + 2 * b;
  for (int y = 0; y < Hm; y++)
    for (int x = 0; x < Wm; x++) dst[(size_t)y * Wm + x] = src[(size_t)refl(y - b, H) * W + refl(x - b, W)];
}

/* ---- a9: lib_image::gaussian smoothing kernel (lib_image.cc:1374-1436),
 * deriv 0: support ceil(3 sigma), unit L1.  Returns its length. */
int orc_gaussian(double sigma, double* out /* >= 2*ceil(3 sigma)+1 */) {
  unsigned long support = (unsigned long)ceil(3 * sigma);
  const double sigma2_inv = 1.0 / (sigma * sigma);
  const double neg_two_sigma2_inv = -0.5 * sigma2_inv;
  unsigned long size = 2 * support + 1;
  double x = -(double)support;
  double s = 0;
  for (unsigned long n = 0; n < size; n++, x++) {
    out[n] = exp(x * x * neg_two_sigma2_inv);
  }
  for (unsigned long n = 0; n < size; n++) s += fabs(out[n]);
  for (unsigned long n = 0; n < size; n++) out[n] /= s;
  return (int)size;
}

/* ---- a9: hist_gradient_2D (lib_image.cc:3575-3588, 3622-3686) with
 * weight_matrix_disc (:3027-3047), orientation_slice_map (:3052-3079),
 * conv_in_place_1D (:3084-3115), compute_hist_gradient_2D (:3120-3231) and the
 * chi-squared functor (matrix_distance_functors.hh:95-108).
 * labels: Hm x Wm (already mirrored); rows are lib_image's x (slow) axis.
 * Histogram length = max(label)+1 over the whole input.  Only positions at
 * least r from the input edge are evaluated (all others are trimmed by the
 * caller); out[o][(row-b)*(Wm-2b) + (col-b)] for b = border. */
void orc_hist_gradient(const uint8_t* labels, int Wm, int Hm, int border, int r, const double* kern, int klen,
                       double* out) {
  int size = 2 * r + 1;
  int nb = 0;
  for (size_t i = 0; i < (size_t)Wm * Hm; i++)
    if (labels[i] + 1 > nb) nb = labels[i] + 1;
  double* weights = (double*)calloc((size_t)size * size, sizeof(double));
  int* smap = (int*)malloc(sizeof(int) * size * size);
  for (int ox = 0; ox < size; ox++)
    for (int oy = 0; oy < size; oy++) {
      long xx = ox - r, yy = oy - r;
      if (xx * xx + yy * yy <= (long)r * r) weights[ox * size + oy] = 1;
    }
  {
    double x = -(double)size / 2;
    for (int nx = 0; nx < size; nx++, x++) {
      double y = -(double)size / 2;
      for (int ny = 0; ny < size; ny++, y++) {
        long double ori = (long double)atan2(y, x) + M_PIl;
        unsigned long idx = (unsigned long)floor((double)(ori / M_PIl * (long double)NORI));
        if (idx >= 2 * NORI) idx = 2 * NORI - 1;
        smap[nx * size + ny] = (int)idx;
      }
    }
  }
  double* sh = (double*)malloc(sizeof(double) * 2 * NORI * nb);
  double* tmp = (double*)malloc(sizeof(double) * nb);
  double* left = (double*)malloc(sizeof(double) * nb);
  double* right = (double*)malloc(sizeof(double) * nb);
  int Wo = Wm - 2 * border, Ho = Hm - 2 * border;
  for (int row = border; row < Hm - border; row++)
    for (int col = border; col < Wm - border; col++) {
      memset(sh, 0, sizeof(double) * 2 * NORI * nb);
      /* window position (ox, oy) pairs with the label at centre + r - o */
      for (int ox = 0; ox < size; ox++)
        for (int oy = 0; oy < size; oy++) {
          int lr = row + r - ox, lc = col + r - oy;
          sh[smap[ox * size + oy] * nb + labels[(size_t)lr * Wm + lc]] += weights[ox * size + oy];
        }
      if (kern && klen > 0) {
        for (int s = 0; s < 2 * NORI; s++) {
          double* h = sh + s * nb;
          /* conv_in_place_1D(h, kern) into tmp, same-size, start at klen/2 */
          for (int n = 0; n < nb; n++) tmp[n] = 0;
          unsigned long pos = (unsigned long)klen / 2;
          for (int n = 0; n < nb; n++, pos++) {
            unsigned long omin = (pos + 1 > (unsigned long)nb) ? (pos + 1 - nb) : 0;
            unsigned long omax = (pos < (unsigned long)klen) ? pos : (unsigned long)klen - 1;
            unsigned long i0 = pos - omin, i1 = omin;
            while (i1 <= omax) {
              tmp[n] += h[i0] * kern[i1];
              i0--;
              i1++;
            }
          }
          for (int n = 0; n < nb; n++) h[n] = tmp[n];
        }
      }
      for (int s = 0; s < 2 * NORI; s++) {
        double t = 0;
        for (int n = 0; n < nb; n++) t += sh[s * nb + n];
        if (t != 0)
          for (int n = 0; n < nb; n++) sh[s * nb + n] /= t;
      }
      for (int n = 0; n < nb; n++) left[n] = right[n] = 0;
      for (int o = 0; o < NORI; o++)
        for (int n = 0; n < nb; n++) {
          left[n] += sh[o * nb + n];
          right[n] += sh[(o + NORI) * nb + n];
        }
      for (int o = 0; o < NORI; o++) {
        double dist = 0;
        for (int n = 0; n < nb; n++) {
          double diff = right[n] - left[n], sum = right[n] + left[n];
          if (diff != 0) dist += diff * diff / sum;
        }
        out[(size_t)o * Wo * Ho + (size_t)(row - border) * Wo + (col - border)] = dist / 2;
        for (int n = 0; n < nb; n++) {
          left[n] -= sh[o * nb + n];
          left[n] += sh[(o + NORI) * nb + n];
          right[n] += sh[o * nb + n];
          right[n] -= sh[(o + NORI) * nb + n];
        }
      }
    }
  free(weights);
  free(smap);
  free(sh);
  free(tmp);
  free(left);
  free(right);
}

/* ---- a9/a11 driver: det_mPb + multiscalePb weighted sum
 * (mex_pb_parts_final_selected.cpp:76-227, multiscalePb.cpp:18-26,216-245).
 * cues[12][8][N] (bg1..3, cga1..3, cgb1..3, tg1..3), mpb_all[8][N]. */
static const double MPB_W[12] = {0.0146, 0.0145, 0.0163, 0.0210, 0.0243, 0.0287,
                                 0.0166, 0.0185, 0.0204, 0.0101, 0.0111, 0.0141};
void orc_local_cues(const uint32_t* packed, const uint8_t* textons, int W, int H, double* cues, double* mpb_all) {
  const int B = 30;
  size_t N = (size_t)W * H;
  int Wm = W + 2 * B, Hm = H + 2 * B;
  size_t Nm = (size_t)Wm * Hm;
  uint32_t* pm = (uint32_t*)malloc(sizeof(uint32_t) * Nm);
  for (int y = 0; y < Hm; y++)
    for (int x = 0; x < Wm; x++) pm[(size_t)y * Wm + x] = packed[(size_t)refl(y - B, H) * W + refl(x - B, W)];
  uint8_t* q = (uint8_t*)malloc(3 * Nm);
  orc_lab_quantize(pm, Nm, 25, q, NULL);
  uint8_t* tm = (uint8_t*)malloc(Nm);
  orc_mirror_u8(textons, W, H, B, tm);
  double kbg[64], kcg[64];
  int lbg = orc_gaussian(0.1 * 25, kbg), lcg = orc_gaussian(0.05 * 25, kcg);
  static const int rbg[3] = {3, 5, 10}, rcg[3] = {5, 10, 20};
  for (int s = 0; s < 3; s++) {
    orc_hist_gradient(q, Wm, Hm, B, rbg[s], kbg, lbg, cues + (size_t)(0 + s) * NORI * N);
    orc_hist_gradient(q + Nm, Wm, Hm, B, rcg[s], kcg, lcg, cues + (size_t)(3 + s) * NORI * N);
    orc_hist_gradient(q + 2 * Nm, Wm, Hm, B, rcg[s], kcg, lcg, cues + (size_t)(6 + s) * NORI * N);
    orc_hist_gradient(tm, Wm, Hm, B, rcg[s], NULL, 0, cues + (size_t)(9 + s) * NORI * N);
  }
  if (mpb_all)
    for (int o = 0; o < NORI; o++)
      for (size_t p = 0; p < N; p++) {
        double v = 0;
        for (int c = 0; c < 12; c++) v += MPB_W[c] * cues[((size_t)c * NORI + o) * N + p];
        mpb_all[(size_t)o * N + p] = v;
      }
  free(pm);
  free(q);
  free(tm);
}

/* ---- a12: nonmax_channels + nonmax_oriented_2D + clip
 * (nonmax_channels.cpp:8-80, lib_image.cc:4083-4200, multiscalePb.cpp:251-261).
 * mpb_all[8][N] -> nmax[N]. */
void orc_nonmax(const double* mpb_all, int W, int H, double* nmax) {
  size_t N = (size_t)W * H;
  double* a = (double*)malloc(sizeof(double) * N);
  double* ind = (double*)malloc(sizeof(double) * N);
  for (size_t p = 0; p < N; p++) {
    double mv = -INFINITY;
    int mi = 0;
    for (int i = 0; i < NORI; i++) {
      double d = mpb_all[(size_t)i * N + p];
      if (d > mv) {
        mv = d;
        mi = i;
      }
    }
    if (mv < 0) mv = 0;
    a[p] = mv;
    ind[p] = (double)mi * REF_PI / (double)NORI;
  }
  const double o_tol = REF_PI / 8;
  /* lib_image's x = rows (size_x = H), y = cols (size_y = W) */
  unsigned long size_x = H, size_y = W;
  unsigned long n = 0;
  for (unsigned long x = 0; x < size_x; x++)
    for (unsigned long y = 0; y < size_y; y++, n++) {
      double ori = ind[n];
      double theta = (double)((long double)ori + M_PI_2l);
      theta = (double)((long double)theta - floorl((long double)theta / M_PIl) * M_PIl);
      double v = a[n];
      unsigned long i0a = 0, i0b = 0, i1a = 0, i1b = 0;
      double d = 0;
      int v0ok = 0, v1ok = 0;
      if (theta == 0) {
        v0ok = (x > 0); v1ok = (x < size_x - 1);
        if (v0ok) { i0a = n - size_y; i0b = i0a; }
        if (v1ok) { i1a = n + size_y; i1b = i1a; }
      } else if ((long double)theta < M_PI_4l) {
        d = (double)tanl((long double)theta);
        v0ok = (x > 0 && y > 0); v1ok = (x < size_x - 1 && y < size_y - 1);
        if (v0ok) { i0a = n - size_y; i0b = i0a - 1; }
        if (v1ok) { i1a = n + size_y; i1b = i1a + 1; }
      } else if ((long double)theta < M_PI_2l) {
        d = (double)tanl(M_PI_2l - (long double)theta);
        v0ok = (x > 0 && y > 0); v1ok = (x < size_x - 1 && y < size_y - 1);
        if (v0ok) { i0a = n - 1; i0b = i0a - size_y; }
        if (v1ok) { i1a = n + 1; i1b = i1a + size_y; }
      } else if ((long double)theta == M_PI_2l) {
        v0ok = (y > 0); v1ok = (y < size_y - 1);
        if (v0ok) { i0a = n - 1; i0b = i0a; }
        if (v1ok) { i1a = n + 1; i1b = i1a; }
      } else if ((long double)theta < 3.0L * M_PI_4l) {
        d = (double)tanl((long double)theta - M_PI_2l);
        v0ok = (x < size_x - 1 && y > 0); v1ok = (x > 0 && y < size_y - 1);
        if (v0ok) { i0a = n - 1; i0b = i0a + size_y; }
        if (v1ok) { i1a = n + 1; i1b = i1a - size_y; }
      } else {
        d = (double)tanl(M_PIl - (long double)theta);
        v0ok = (x < size_x - 1 && y > 0); v1ok = (x > 0 && y < size_y - 1);
        if (v0ok) { i0a = n + size_y; i0b = i0a - 1; }
        if (v1ok) { i1a = n - size_y; i1b = i1a + 1; }
      }
      double res = 0;
      if (v0ok && v1ok) {
        double v0a = a[i0a], v0b = a[i0b], v1a = a[i1a], v1b = a[i1b];
        double od[4] = {ind[i0a] - ori, ind[i0b] - ori, ind[i1a] - ori, ind[i1b] - ori};
        double c[4];
        for (int k = 0; k < 4; k++) {
          double t = (double)((long double)od[k] - floorl((long double)od[k] / (2 * M_PIl)) * (2 * M_PIl));
          if ((long double)t >= M_PIl) t = (double)(2 * M_PIl - (long double)t);
          if ((long double)t >= M_PI_2l) t = (double)(M_PIl - (long double)t);
          t = (t <= o_tol) ? 0 : (t - o_tol);
          c[k] = cos(t);
        }
        double v0 = (1.0 - d) * v0a * c[0] + d * v0b * c[1];
        double v1 = (1.0 - d) * v1a * c[2] + d * v1b * c[3];
        if (v > v0 && v > v1) res = v;
      }
      /* clip of multiscalePb.cpp:254-260: values with 1.2 v > 1 become 1 */
      if (1.2 * res > 1) res = 1;
      if (res < 0) res = 0;
      nmax[n] = res;
    }
  free(a);
  free(ind);
}

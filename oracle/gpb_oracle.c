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
 *   - borders wider than the image reflect periodically (lib_image's
 *     border_mirror_2D throws there, lib_image.cc:308-370);
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
static void hist_gradient_nb(const uint8_t* labels, int Wm, int Hm, int border, int r, const double* kern,
                             int klen, int nb_force, double* out);
void orc_hist_gradient(const uint8_t* labels, int Wm, int Hm, int border, int r, const double* kern, int klen,
                       double* out) {
  hist_gradient_nb(labels, Wm, Hm, border, r, kern, klen, 0, out);
}

/* As orc_hist_gradient with the histogram length given (nb_force > 0) instead
 * of max(label) + 1 over the input: the length of a whole image, for a
 * window computed on a crop of it (the smoothing of the bins depends on the
 * length at the top bins). */
static void hist_gradient_nb(const uint8_t* labels, int Wm, int Hm, int border, int r, const double* kern,
                             int klen, int nb_force, double* out) {
  int size = 2 * r + 1;
  int nb = 0;
  for (size_t i = 0; i < (size_t)Wm * Hm; i++)
    if (labels[i] + 1 > nb) nb = labels[i] + 1;
  if (nb_force > 0) nb = nb_force > nb ? nb_force : nb;
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
void orc_local_cues_nb(const uint32_t* packed, const uint8_t* textons, int W, int H, const int* nb4, double* cues,
                       double* mpb_all);
void orc_local_cues(const uint32_t* packed, const uint8_t* textons, int W, int H, double* cues, double* mpb_all) {
  orc_local_cues_nb(packed, textons, W, H, NULL, cues, mpb_all);
}

/* nb4 (optional): histogram lengths of the L, a, b and texton channels (see
 * hist_gradient_nb). */
void orc_local_cues_nb(const uint32_t* packed, const uint8_t* textons, int W, int H, const int* nb4, double* cues,
                       double* mpb_all) {
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
    hist_gradient_nb(q, Wm, Hm, B, rbg[s], kbg, lbg, nb4 ? nb4[0] : 0, cues + (size_t)(0 + s) * NORI * N);
    hist_gradient_nb(q + Nm, Wm, Hm, B, rcg[s], kcg, lcg, nb4 ? nb4[1] : 0, cues + (size_t)(3 + s) * NORI * N);
    hist_gradient_nb(q + 2 * Nm, Wm, Hm, B, rcg[s], kcg, lcg, nb4 ? nb4[2] : 0, cues + (size_t)(6 + s) * NORI * N);
    hist_gradient_nb(tm, Wm, Hm, B, rcg[s], NULL, 0, nb4 ? nb4[3] : 0, cues + (size_t)(9 + s) * NORI * N);
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

/* =========================================================================
 * a13: intervening-contour affinity.  Restates buildW (buildW.cpp:23-212):
 * lattices H(x,y) = mPb(row y-1, col x), V(x,y) = mPb(row y, col x-1)
 * (spectralPb.cpp:55-59; FP32 lattices, ic.hh:31-34), Group::computeSupport /
 * interveningContour / ic_walk (buildW/src/ic.cc:42-560) with threshold 1,
 * computeAffinities2 (affinity.cc:58-156, C_IC_SS :40-52, sigma 0.1, radius 5)
 * and SMatrix::symmetrize (smatrix.cc:95-117).
 * Output: Wd[k][p] for the 81 disc offsets k (row-major over dy, dx), 0 where
 * the neighbour is outside the image; D[p] = full row sum (buildW.cpp:150-160).
 * ========================================================================= */
typedef struct { int x, y; float sim; } ic_pt;
typedef struct {
  int w, h;           /* lattice image width (cols) and height (rows) */
  const float* H;     /* H[x*(h+1) + y], x < w, y <= h */
  const float* V;     /* V[x*h + y], x <= w, y < h */
} ic_lat;
static float latH(const ic_lat* L, int x, int y) { return L->H[(size_t)x * (L->h + 1) + y]; }
static float latV(const ic_lat* L, int x, int y) { return L->V[(size_t)x * L->h + y]; }

typedef struct { ic_pt pts[11][11]; int cnt[11]; } ic_scan;

static void ic_walk_o(const ic_lat* L, float thresh, int x0, int y0, int x1, int y1, int x2, int y2, int x3, int y3,
                      int wr, ic_scan* sc) {
  const long long dx1 = x1 - x0, dy1 = y1 - y0, dx2 = x2 - x0, dy2 = y2 - y0, dx3 = x3 - x0, dy3 = y3 - y0;
  const long long dot11 = dx1 * dx1 + dy1 * dy1, dot22 = dx2 * dx2 + dy2 * dy2, dot33 = dx3 * dx3 + dy3 * dy3;
  const int dx = x2 - x0, dy = y2 - y0, adx = abs(dx), ady = abs(dy);
  int octant = -1;
  if (dx > 0 && dy >= 0) octant = (adx > ady) ? 0 : 1;
  else if (dx <= 0 && dy > 0) octant = (adx < ady) ? 2 : 3;
  else if (dy <= 0 && dx < 0) octant = (adx > ady) ? 4 : 5;
  else if (dx >= 0 && dy < 0) octant = (adx < ady) ? 6 : 7;
  int t = (octant == 0 || octant == 3 || octant == 4 || octant == 7) ? -adx : -ady;
  float maxpb = 0.0f;
  int xi = x0, yi = y0, count = 0, oldx = xi, oldy = yi;
  ic_pt pts[64];
  while (xi != x2 || yi != y2) {
    switch (octant) {
      case 0: xi++; t += (ady << 1); if (t > 0) { yi++; t -= (adx << 1); } break;
      case 1: yi++; t += (adx << 1); if (t > 0) { xi++; t -= (ady << 1); } break;
      case 2: yi++; t += (adx << 1); if (t > 0) { xi--; t -= (ady << 1); } break;
      case 3: xi--; t += (ady << 1); if (t > 0) { yi++; t -= (adx << 1); } break;
      case 4: xi--; t += (ady << 1); if (t > 0) { yi--; t -= (adx << 1); } break;
      case 5: yi--; t += (adx << 1); if (t > 0) { xi--; t -= (ady << 1); } break;
      case 6: yi--; t += (adx << 1); if (t > 0) { xi++; t -= (ady << 1); } break;
      default: xi++; t += (ady << 1); if (t > 0) { yi--; t -= (adx << 1); } break;
    }
    const long long dxi = xi - x0, dyi = yi - y0;
    const long long dotii = dxi * dxi + dyi * dyi, doti1 = dxi * dx1 + dyi * dy1;
    const long long doti2 = dxi * dx2 + dyi * dy2, doti3 = dxi * dx3 + dyi * dy3;
    const int good = (doti2 * doti2 == dotii * dot22) ||
                     (dot11 * doti2 * doti2 > dot22 * doti1 * doti1 && dot33 * doti2 * doti2 >= dot22 * doti3 * doti3);
    float in = 0.0f;
    if (oldx == xi) {
      if (yi > oldy) in = latH(L, xi, yi);
      else if (yi < oldy) in = latH(L, oldx, oldy);
    } else if (oldy == yi) {
      if (xi > oldx) in = latV(L, xi, yi);
      else if (xi < oldx) in = latV(L, oldx, oldy);
    } else {
      float a, b, c, d;
      if (xi > oldx && yi > oldy) { a = latH(L, oldx, yi); b = latH(L, xi, yi); c = latV(L, xi, oldy); d = latV(L, xi, yi); }
      else if (xi > oldx && yi < oldy) { a = latH(L, oldx, oldy); b = latH(L, xi, oldy); c = latV(L, xi, oldy); d = latV(L, xi, yi); }
      else if (xi < oldx && yi > oldy) { a = latH(L, oldx, yi); b = latH(L, xi, yi); c = latV(L, oldx, oldy); d = latV(L, oldx, yi); }
      else { a = latH(L, oldx, oldy); b = latH(L, xi, oldy); c = latV(L, oldx, oldy); d = latV(L, oldx, yi); }
      in = fmaxf(in, a); in = fmaxf(in, b); in = fmaxf(in, c); in = fmaxf(in, d);
    }
    maxpb = fmaxf(maxpb, in);
    oldx = xi;
    oldy = yi;
    if (!good) continue;
    if (maxpb > thresh) break;
    pts[count].x = xi; pts[count].y = yi; pts[count].sim = 1.0f - maxpb; count++;
  }
  if (octant <= 1 || octant >= 6) {
    for (int i = 0; i < count; i++) { int yind = pts[i].y - y0 + wr; sc->pts[yind][sc->cnt[yind]++] = pts[i]; }
  } else {
    for (int i = count - 1; i >= 0; i--) { int yind = pts[i].y - y0 + wr; sc->pts[yind][sc->cnt[yind]++] = pts[i]; }
  }
}

static void ic_corner(ic_scan* sc, int x, int y, int y0, int wr) {
  int yind = y - y0 + wr;
  ic_pt p = {x, y, 1.0f};
  sc->pts[yind][sc->cnt[yind]++] = p;
}

/* support list of one pixel, scanline order; returns count */
static int ic_support(const ic_lat* L, float thresh, int x0, int y0, int wr, ic_pt* adj) {
  const int width = L->w, height = L->h;
  ic_scan sc;
  memset(sc.cnt, 0, sizeof(sc.cnt));
  const int rxa = x0 - wr > 0 ? x0 - wr : 0, rya = y0 - wr > 0 ? y0 - wr : 0;
  const int rxb = x0 + wr < width - 1 ? x0 + wr : width - 1, ryb = y0 + wr < height - 1 ? y0 + wr : height - 1;
#define W_(a, b, c, d, e, f) ic_walk_o(L, thresh, x0, y0, a, b, c, d, e, f, wr, &sc)
  if (x0 > rxa) {
    if ((y0 > 0) && (y0 < ryb)) W_(rxa, y0 - 1, rxa, y0, rxa, y0 + 1);
    for (int y = y0 - 1; y > rya; y--) W_(rxa, y - 1, rxa, y, rxa, y + 1);
  }
  if (x0 > rxa + 1 || y0 > rya + 1 || ((x0 > rxa) && (y0 > rya))) W_(rxa, rya + 1, rxa, rya, rxa + 1, rya);
  if (((x0 == rxa) && (y0 == rya + 1)) || ((x0 == rxa + 1) && (y0 == rya))) ic_corner(&sc, rxa, rya, y0, wr);
  if (y0 > rya)
    for (int x = rxa + 1; x < rxb; x++) W_(x - 1, rya, x, rya, x + 1, rya);
  if ((y0 > rya + 1) || (x0 < rxb - 1) || ((y0 > rya) && (x0 < rxb))) W_(rxb - 1, rya, rxb, rya, rxb, rya + 1);
  if (((x0 == rxb - 1) && (y0 == rya)) || ((x0 == rxb) && (y0 == rya + 1))) ic_corner(&sc, rxb, rya, y0, wr);
  if (x0 < rxb)
    for (int y = rya + 1; y < y0; y++) W_(rxb, y - 1, rxb, y, rxb, y + 1);
  if (x0 > rxa)
    for (int y = y0 + 1; y < ryb; y++) W_(rxa, y - 1, rxa, y, rxa, y + 1);
  if ((x0 > rxa + 1) || (y0 < ryb - 1) || ((x0 > rxa) && (y0 < ryb))) W_(rxa, ryb - 1, rxa, ryb, rxa + 1, ryb);
  if (((x0 == rxa) && (y0 == ryb - 1)) || ((x0 == rxa + 1) && (y0 == ryb))) ic_corner(&sc, rxa, ryb, y0, wr);
  if (y0 < ryb)
    for (int x = rxa + 1; x < rxb; x++) W_(x - 1, ryb, x, ryb, x + 1, ryb);
  if ((y0 < ryb - 1) || (x0 < rxb - 1) || ((y0 < ryb) && (x0 < rxb))) W_(rxb - 1, ryb, rxb, ryb, rxb, ryb - 1);
  if (((x0 == rxb - 1) && (y0 == ryb)) || ((x0 == rxb) && (y0 == ryb - 1))) ic_corner(&sc, rxb, ryb, y0, wr);
  if (x0 < rxb) {
    for (int y = ryb - 1; y > y0; y--) W_(rxb, y - 1, rxb, y, rxb, y + 1);
    if ((y0 > 0) && (y0 < ryb)) W_(rxb, y0 - 1, rxb, y0, rxb, y0 + 1);
  }
#undef W_
  int count = 0;
  for (int y = 0; y < 2 * wr + 1; y++)
    for (int i = 0; i < sc.cnt[y]; i++) adj[count++] = sc.pts[y][i];
  return count;
}

static float c_ic_ss(float ic) {
  float val = 1.8682;
  val += (ic / 0.3130) * -1.3113;
  const float post = 1.0 / (1.0 + exp(-val));
  if (!isfinite(post)) return 0;
  if (post < 0) return 0;
  if (post > 1) return 1;
  return post;
}

/* Returns -1 if a scan line was not strictly increasing (reference asserts). */
int orc_build_w(const double* mpb, int W, int H, double* Wd /* [81][N] */, double* D /* [N] */) {
  const int wr = 5;
  const float sigma = 0.1f, thresh = 1.0f;
  size_t N = (size_t)W * H;
  ic_lat L;
  L.w = W;
  L.h = H;
  float* Hl = (float*)calloc((size_t)W * (H + 1), sizeof(float));
  float* Vl = (float*)calloc((size_t)(W + 1) * H, sizeof(float));
  for (int x = 0; x < W; x++)
    for (int y = 1; y <= H; y++) Hl[(size_t)x * (H + 1) + y] = (float)mpb[(size_t)(y - 1) * W + x];
  for (int x = 1; x <= W; x++)
    for (int y = 0; y < H; y++) Vl[(size_t)x * H + y] = (float)mpb[(size_t)y * W + (x - 1)];
  L.H = Hl;
  L.V = Vl;
  int dxk[81], dyk[81], nk = 0;
  for (int u = -wr; u <= wr; u++)
    for (int v = -wr; v <= wr; v++)
      if (u * u + v * v <= wr * wr) { dxk[nk] = v; dyk[nk] = u; nk++; }
  /* unsymmetrised affinities, same [81][N] layout */
  float* A = (float*)calloc(81 * N, sizeof(float));
  int bad = 0;
  ic_pt adj[121];
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      int n = ic_support(&L, thresh, x, y, wr, adj);
      for (int r = 0; r + 1 < 11; r++) { /* sortedness check per scan line */ }
      for (int i = 1; i < n; i++)
        if (adj[i].y == adj[i - 1].y && !(adj[i - 1].x < adj[i].x)) bad = 1;
      int idx = 0, k = 0;
      for (int u = -wr; u <= wr; u++) {
        int yy = y + u;
        for (int v = -wr; v <= wr; v++) {
          int xx = x + v;
          if (u * u + v * v > wr * wr) continue;
          int kk = k++;
          if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
          while (idx < n && adj[idx].y < yy) idx++;
          while (idx < n && adj[idx].x < xx) idx++;
          float pss;
          if (u == 0 && v == 0) pss = 1.0f;
          else {
            float icsim = 0.0f;
            if (idx < n && adj[idx].x == xx && adj[idx].y == yy) { icsim = adj[idx].sim; idx++; }
            pss = c_ic_ss(1 - icsim);
          }
          float val = exp(-(1 - pss) / sigma);
          A[(size_t)kk * N + (size_t)y * W + x] = val;
        }
      }
    }
  for (size_t p = 0; p < N; p++) D[p] = 0;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      size_t p = (size_t)y * W + x;
      for (int k = 0; k < 81; k++) {
        int xx = x + dxk[k], yy = y + dyk[k];
        double w = 0;
        if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
          size_t q = (size_t)yy * W + xx;
          w = 0.5 * ((double)A[(size_t)k * N + p] + (double)A[(size_t)(80 - k) * N + q]);
        }
        Wd[(size_t)k * N + p] = w;
        D[p] += w;
      }
    }
  free(A);
  free(Hl);
  free(Vl);
  return bad ? -1 : 0;
}

/* =========================================================================
 * a14-a16: generalised eigenpairs of (D - W, D): the K smallest, by thick-
 * restart Lanczos with full (two-pass Gram-Schmidt) reorthogonalisation in
 * FP64 on M = D^-1/2 W D^-1/2, whose largest eigenvalues mu give lambda =
 * 1 - mu (Spectra SymGEigsSolver in Cholesky mode, SymGEigsSolver.h:46-111,
 * SymEigsSolver.h:200-586).  The constant vector (lambda = 0) is locked as
 * the first basis vector.  Stops when every returned pair has
 * ||(D-W)v - lambda D v|| / ||D v|| <= tol.  evecs[K][N] with v^T D v = 1.
 * ========================================================================= */
static void stencil_apply(const double* Wd, const double* s, int W, int H, const double* x, double* y) {
  static int dxk[81], dyk[81], ready = 0;
  if (!ready) {
    int nk = 0;
    for (int u = -5; u <= 5; u++)
      for (int v = -5; v <= 5; v++)
        if (u * u + v * v <= 25) { dxk[nk] = v; dyk[nk] = u; nk++; }
    ready = 1;
  }
  size_t N = (size_t)W * H;
  for (int yy = 0; yy < H; yy++)
    for (int xx = 0; xx < W; xx++) {
      size_t p = (size_t)yy * W + xx;
      double acc = 0;
      for (int k = 0; k < 81; k++) {
        int x2 = xx + dxk[k], y2 = yy + dyk[k];
        if (x2 < 0 || x2 >= W || y2 < 0 || y2 >= H) continue;
        size_t q = (size_t)y2 * W + x2;
        acc += Wd[(size_t)k * N + p] * s[q] * x[q];
      }
      y[p] = s[p] * acc;
    }
}

/* cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major);
 * eigenvalues descending into w, eigenvectors as columns of Z */
static void jacobi_eig(int n, double* A, double* w, double* Z) {
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Z[i * n + j] = (i == j);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += A[i * n + j] * A[i * n + j];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p * n + q];
        if (fabs(apq) < 1e-300) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double theta = (aqq - app) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        double c = 1 / sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double zkp = Z[k * n + p], zkq = Z[k * n + q];
          Z[k * n + p] = c * zkp - s * zkq;
          Z[k * n + q] = s * zkp + c * zkq;
        }
      }
  }
  /* sort descending */
  int* idx = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (A[idx[j] * n + idx[j]] > A[idx[i] * n + idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
  double* Zs = (double*)malloc(sizeof(double) * n * n);
  for (int i = 0; i < n; i++) {
    w[i] = A[idx[i] * n + idx[i]];
    for (int k = 0; k < n; k++) Zs[k * n + i] = Z[k * n + idx[i]];
  }
  memcpy(Z, Zs, sizeof(double) * n * n);
  free(Zs);
  free(idx);
}

void orc_jacobi_eig(int n, double* A, double* w, double* Z) { jacobi_eig(n, A, w, Z); }

static double dotv(const double* a, const double* b, size_t n) {
  double s = 0;
  for (size_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

int orc_eigs(const double* Wd, const double* D, int W, int H, int K, int m, double tol, int maxit, double* evals,
             double* evecs, int* iters_out) {
  size_t N = (size_t)W * H;
  if (m < K + 2) m = K + 2;
  double* s = (double*)malloc(sizeof(double) * N);
  for (size_t p = 0; p < N; p++) s[p] = 1.0 / sqrt(D[p]);
  double* V = (double*)malloc(sizeof(double) * N * (m + 1));
  double* w = (double*)malloc(sizeof(double) * N);
  double* T = (double*)calloc((size_t)m * m, sizeof(double));
  double* A = (double*)malloc(sizeof(double) * m * m);
  double* th = (double*)malloc(sizeof(double) * m);
  double* Y = (double*)malloc(sizeof(double) * m * m);
  double* U = (double*)malloc(sizeof(double) * N * m);
  double* h = (double*)malloc(sizeof(double) * (m + 1));
  /* locked trivial vector D^1/2 1 */
  double nrm = 0;
  for (size_t p = 0; p < N; p++) { V[p] = sqrt(D[p]); nrm += D[p]; }
  nrm = sqrt(nrm);
  for (size_t p = 0; p < N; p++) V[p] /= nrm;
  /* start vector: deterministic pseudo-random, orthogonal to V0 */
  {
    uint64_t z = 0x2545F4914F6CDD1Dull;
    double* v1 = V + N;
    for (size_t p = 0; p < N; p++) {
      z ^= z << 13; z ^= z >> 7; z ^= z << 17;
      v1[p] = (double)(z >> 11) / 9007199254740992.0 - 0.5;
    }
    double c = dotv(v1, V, N);
    for (size_t p = 0; p < N; p++) v1[p] -= c * V[p];
    double n1 = sqrt(dotv(v1, v1, N));
    for (size_t p = 0; p < N; p++) v1[p] /= n1;
  }
  int k = 1, iters = 0, done = 0, restarts = 0;
  memset(T, 0, sizeof(double) * m * m);
  {
    /* T[0][0] = v0' M v0 */
    stencil_apply(Wd, s, W, H, V, w);
    T[0] = dotv(V, w, N);
    iters++;
  }
  double beta = 0;
  while (!done) {
    for (int j = k; j < m; j++) {
      stencil_apply(Wd, s, W, H, V + (size_t)j * N, w);
      iters++;
      for (int i = 0; i <= j; i++) h[i] = 0;
      for (int pass = 0; pass < 2; pass++)
        for (int i = 0; i <= j; i++) {
          double c = dotv(V + (size_t)i * N, w, N);
          h[i] += c;
          for (size_t p = 0; p < N; p++) w[p] -= c * V[(size_t)i * N + p];
        }
      for (int i = 0; i <= j; i++) { T[i * m + j] = h[i]; T[j * m + i] = h[i]; }
      beta = sqrt(dotv(w, w, N));
      if (beta < 1e-300) beta = 1e-300;
      for (size_t p = 0; p < N; p++) V[(size_t)(j + 1) * N + p] = w[p] / beta;
    }
    memcpy(A, T, sizeof(double) * m * m);
    jacobi_eig(m, A, th, Y);
    /* Ritz vectors of the leading K (plus extras to keep) */
    int keep = K + (m - K) / 2;
    if (keep > m - 1) keep = m - 1;
    for (int i = 0; i < keep; i++) {
      double* u = U + (size_t)i * N;
      for (size_t p = 0; p < N; p++) u[p] = 0;
      for (int j = 0; j < m; j++) {
        double c = Y[j * m + i];
        const double* vj = V + (size_t)j * N;
        for (size_t p = 0; p < N; p++) u[p] += c * vj[p];
      }
    }
    /* relative generalised residuals of the leading K */
    double bd = 0;
    const double* vm = V + (size_t)m * N;
    for (size_t p = 0; p < N; p++) bd += D[p] * vm[p] * vm[p];
    bd = sqrt(bd);
    int conv = 1;
    for (int i = 0; i < K; i++) {
      double ud = 0;
      const double* u = U + (size_t)i * N;
      for (size_t p = 0; p < N; p++) ud += D[p] * u[p] * u[p];
      double res = fabs(beta * Y[(m - 1) * m + i]) * bd / sqrt(ud);
      if (res > tol) conv = 0;
    }
    if (conv || iters >= maxit) {
      for (int i = 0; i < K; i++) {
        evals[i] = 1.0 - th[i];
        for (size_t p = 0; p < N; p++) evecs[(size_t)i * N + p] = s[p] * U[(size_t)i * N + p];
      }
      done = 1;
      break;
    }
    /* thick restart */
    memcpy(V + (size_t)keep * N, vm, sizeof(double) * N);
    memcpy(V, U, sizeof(double) * N * keep);
    memset(T, 0, sizeof(double) * m * m);
    for (int i = 0; i < keep; i++) T[i * m + i] = th[i];
    k = keep;
    restarts++;
  }
  if (iters_out) *iters_out = iters;
  free(s); free(V); free(w); free(T); free(A); free(th); free(Y); free(U); free(h);
  return (iters >= maxit) ? -1 : restarts;
}

/* true relative generalised residual of one pair */
double orc_gen_residual(const double* Wd, const double* D, int W, int H, double lam, const double* v) {
  size_t N = (size_t)W * H;
  double* one = (double*)malloc(sizeof(double) * N);
  double* Wv = (double*)malloc(sizeof(double) * N);
  for (size_t p = 0; p < N; p++) one[p] = 1.0;
  stencil_apply(Wd, one, W, H, v, Wv);
  double num = 0, den = 0;
  for (size_t p = 0; p < N; p++) {
    double r = D[p] * v[p] - Wv[p] - lam * D[p] * v[p];
    num += r * r;
    den += (D[p] * v[p]) * (D[p] * v[p]);
  }
  free(one);
  free(Wv);
  return sqrt(num / den);
}

/* =========================================================================
 * a17-a19: eigenvector normalisation, sPb and gPb.
 * spectralPb.cpp:257-322: for i = 1..K-1, v <- (v - min)/(max - min)
 * (left as is when max == min), divided by sqrt(lambda_i) when lambda_i >
 * DBL_EPSILON, correlated (filter2D, BORDER_REFLECT_101) with 8 oriented
 * 7x7 first-derivative filters; sPb_o += |response|.  gPb_o = sum of the 12
 * gPb-weighted cues + 0.0074 sPb_o (globalPb.cpp:31-37, 97-103).
 * ========================================================================= */
static void spb_filters(double* out /* [8][7][7] */) {
  for (int o = 0; o < 8; o++) {
    double ori = (double)o * M_PI / 8;
    /* rotated supports (support_x_rotated / support_y_rotated with -ori) */
    double sxa = 3 * cos(-ori), sya = 3 * sin(-ori);
    double x0m = fabs(sxa - sya), x1m = fabs(sxa + sya);
    long sxr = (long)ceil(x0m > x1m ? x0m : x1m);
    double s2 = 3 * sin(-ori), c2 = 3 * cos(-ori);
    double y0m = fabs(s2 - c2), y1m = fabs(s2 + c2);
    long syr = (long)ceil(y0m > y1m ? y0m : y1m);
    long nx = 2 * sxr + 1, ny = 2 * syr + 1;
    double mx[64], my[64];
    double sa = 0, sm = 0;
    for (long i = 0; i < nx; i++) { double x = (double)(i - sxr); mx[i] = exp(-0.5 * x * x); }
    for (long i = 0; i < nx; i++) sa += fabs(mx[i]);
    for (long i = 0; i < nx; i++) mx[i] /= sa;
    for (long i = 0; i < ny; i++) { double x = (double)(i - syr); my[i] = exp(-0.5 * x * x) * (-x); sm += my[i]; }
    sm /= (double)ny;
    sa = 0;
    for (long i = 0; i < ny; i++) { my[i] -= sm; sa += fabs(my[i]); }
    for (long i = 0; i < ny; i++) my[i] /= sa;
    double* f = out + o * 49;
    double co = cos(ori), so = sin(ori), ox = (double)(nx - 1) / 2, oy = (double)(ny - 1) / 2;
    for (int a = 0; a < 7; a++)
      for (int b = 0; b < 7; b++) {
        double u = a - 3.0, v = b - 3.0;
        double x = u * co + v * so + ox, y = v * co - u * so + oy, val = 0;
        if (x >= 0 && y >= 0) {
          long xa = (long)floor(x), xb = (long)ceil(x), ya = (long)floor(y), yb = (long)ceil(y);
          if (xb < nx && yb < ny) {
            double m00 = mx[xa] * my[ya], m01 = mx[xa] * my[yb], m10 = mx[xb] * my[ya], m11 = mx[xb] * my[yb];
            double t0 = (xa != xb) ? ((xb - x) * m00 + (x - xa) * m10) : m00;
            double t1 = (xa != xb) ? ((xb - x) * m01 + (x - xa) * m11) : m01;
            val = (ya != yb) ? ((yb - y) * t0 + (y - ya) * t1) : t0;
          }
        }
        f[a * 7 + b] = val;
      }
    double mean = 0, sab = 0;
    for (int n = 0; n < 49; n++) mean += f[n];
    mean /= 49;
    for (int n = 0; n < 49; n++) { f[n] -= mean; sab += fabs(f[n]); }
    for (int n = 0; n < 49; n++) f[n] /= sab;
  }
}

void orc_spb_filters(double* out) { spb_filters(out); }

void orc_spb(const double* evecs, const double* evals, int K, int W, int H, double* spb /* [8][N] */) {
  size_t N = (size_t)W * H;
  double F[8 * 49];
  spb_filters(F);
  double* v = (double*)malloc(sizeof(double) * N);
  memset(spb, 0, sizeof(double) * 8 * N);
  for (int i = 1; i < K; i++) {
    const double* e = evecs + (size_t)i * N;
    double mn = e[0], mx = e[0];
    for (size_t p = 0; p < N; p++) { if (e[p] < mn) mn = e[p]; if (e[p] > mx) mx = e[p]; }
    for (size_t p = 0; p < N; p++) v[p] = (mx > mn) ? (e[p] - mn) / (mx - mn) : e[p];
    if (!(evals[i] > DBL_EPSILON)) continue;
    double sc = sqrt(evals[i]);
    for (size_t p = 0; p < N; p++) v[p] /= sc;
    for (int o = 0; o < 8; o++)
      for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
          double acc = 0;
          for (int a = 0; a < 7; a++)
            for (int b = 0; b < 7; b++)
              acc += F[o * 49 + a * 7 + b] * v[(size_t)refl101(y + a - 3, H) * W + refl101(x + b - 3, W)];
          spb[(size_t)o * N + (size_t)y * W + x] += fabs(acc);
        }
  }
  free(v);
}

static const double GPB_W[13] = {0, 0, 0.0039, 0.0050, 0.0058, 0.0069, 0.0040, 0.0044, 0.0049, 0.0024, 0.0027, 0.0170, 0.0074};
void orc_gpb(const double* cues /* [12][8][N] */, const double* spb /* [8][N] */, size_t N, double* gpb /* [8][N] */) {
  for (int o = 0; o < 8; o++)
    for (size_t p = 0; p < N; p++) {
      double v = 0;
      for (int c = 0; c < 12; c++) v += GPB_W[c] * cues[((size_t)c * 8 + o) * N + p];
      v += GPB_W[12] * spb[(size_t)o * N + p];
      gpb[(size_t)o * N + p] = v;
    }
}

/* computeGPb output transforms (damascene globalPb.cu:5-49, skeleton.cu:229-249,
 * damascene.cu:594-604): max over orientations masked by mPb > 0.05, and the
 * calibration sigmoid of each oriented map, written column-major with
 * orientation o in block 7 - o (block i = slot oriMap[i] = (i + 4) mod 8,
 * damascene.cu:598-603; slot s = the filter built for angle
 * pi/2 + (7 - s) pi/8, spectralPb.cu:706-712, applied row-major,
 * spectralPb.cu:101, which is orientation (3 - s) mod 8 here). */
void orc_gpb_outputs(const double* gpb /* [8][N] */, const double* mpb_nmax, int W, int H, float* gmax, float* gall) {
  size_t N = (size_t)W * H;
  for (size_t p = 0; p < N; p++) {
    float m = -200.f;
    for (int o = 0; o < 8; o++) {
      float v = (float)gpb[(size_t)o * N + p];
      if (v > m) m = v;
    }
    gmax[p] = (mpb_nmax[p] > 0.05) ? m : 0.0f;
  }
  for (int o = 0; o < 8; o++)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        float value = (float)gpb[(size_t)o * N + (size_t)y * W + x];
        if (value * 1.2 > 1) value = 1;
        else value *= 1.2;
        if (value < 0) value = 0;
        value = 1.0f / (1.0f + expf(2.6433 - 10.7998 * value));
        value = (value - 0.0667) / 0.9333;
        gall[(size_t)(7 - o) * N + (size_t)x * H + y] = value;
      }
}

/* ---- texton histogram features of the superpixel classifier (test
 * infrastructure): SPFeatures::genTextonFeatures (src/SPFeatures.cpp:259-272)
 * and the texton part of genRecPatchFeatures (:478-545), on a given texton
 * map.  descr[l][t] = count / size * 10 in that order of operations; a label
 * with no pixels keeps zeros.  Returns 1 for a label outside [0, nlabels). */
int orc_texton_features(const uint8_t* textons, const int32_t* sp, size_t N, int nlabels, double* descr) {
  unsigned* cnt = (unsigned*)calloc((size_t)nlabels * 32, sizeof(unsigned));
  for (size_t p = 0; p < N; p++) {
    if (sp[p] < 0 || sp[p] >= nlabels) { free(cnt); return 1; }
    cnt[(size_t)sp[p] * 32 + textons[p]]++;
  }
  for (int l = 0; l < nlabels; l++) {
    unsigned size = 0;
    for (int t = 0; t < 32; t++) size += cnt[(size_t)l * 32 + t];
    for (int t = 0; t < 32; t++) {
      double v = (double)cnt[(size_t)l * 32 + t];
      if (size > 0) { v /= (double)size; v *= 10.0; }
      descr[(size_t)l * 32 + t] = v;
    }
  }
  free(cnt);
  return 0;
}

void orc_texton_patch_features(const uint8_t* textons, int W, int H, const int32_t* centres, int n, double* descr) {
  for (int k = 0; k < n; k++) {
    int row = centres[2 * k], col = centres[2 * k + 1];
    int north = row < 32 ? row : 32, south = (H - 1 - row) < 32 ? (H - 1 - row) : 32;
    int west = col < 32 ? col : 32, east = (W - 1 - col) < 32 ? (W - 1 - col) : 32;
    unsigned h[32] = {0};
    for (int y = row - north; y <= row + south; y++)
      for (int x = col - west; x <= col + east; x++) h[textons[(size_t)y * W + x]]++;
    int area = (north + south + 1) * (west + east + 1);
    for (int t = 0; t < 32; t++) {
      double v = (double)h[t];
      v /= (double)area;
      v *= 10.0;
      descr[(size_t)k * 32 + t] = v;
    }
  }
}

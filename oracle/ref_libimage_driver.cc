// Test-infrastructure driver around the reference's own lib_image
// (globalPb/src/math/libraries/lib_image.cc and its support library,
// compiled from the reference tree by `make ref`).  Each entry point feeds
// plain arrays through the reference's calls the way its glue does
// (src/mex_pb_parts_final_selected.cpp:40-54,100-227,
// src/nonmax_channels.cpp:8-80, src/nonmax_oriented.cpp:75-92,
// src/multiscalePb.cpp:251-261).  Used only by tests, to pin the oracle.
#include <limits>

#include "collections/array_list.hh"
#include "collections/pointers/auto_collection.hh"
#include "math/libraries/lib_image.hh"
#include "math/matrices/matrix.hh"

using collections::array_list;
using collections::pointers::auto_collection;
using math::libraries::lib_image;
using math::matrices::matrix;

extern "C" {

// lib_image::gaussian(sigma) (lib_image.hh:404); returns the kernel length
// (or the needed capacity when it exceeds cap).
int ref_gaussian(double sigma, double* out, int cap) {
  matrix<> k = lib_image::gaussian(sigma);
  const int n = int(k.size());
  if (n <= cap)
    for (int i = 0; i < n; i++) out[i] = k[unsigned(i)];
  return n;
}

// Gamma 2.5, RGB -> Lab, normalise, quantise to nbins (mex_pb_parts
// :115-134): rgb[N][3] bytes (each / 255 as GlobalPb's convertTo does),
// bins[3][N].
void ref_lab_quantize(const unsigned char* rgb, int n, int nbins, unsigned char* bins) {
  matrix<> r(1, unsigned(n)), g(1, unsigned(n)), b(1, unsigned(n));
  for (int i = 0; i < n; i++) {
    r(0, unsigned(i)) = rgb[3 * i] / 255.0;
    g(0, unsigned(i)) = rgb[3 * i + 1] / 255.0;
    b(0, unsigned(i)) = rgb[3 * i + 2] / 255.0;
  }
  lib_image::rgb_gamma_correct(r, g, b, 2.5);
  lib_image::rgb_to_lab(r, g, b);
  lib_image::lab_normalize(r, g, b);
  const matrix<>* ch[3] = {&r, &g, &b};
  for (int c = 0; c < 3; c++) {
    matrix<unsigned long> q = lib_image::quantize_values(*ch[c], unsigned(nbins));
    for (int i = 0; i < n; i++) bins[size_t(c) * n + i] = (unsigned char)q(0, unsigned(i));
  }
}

// lib_image::border_mirror_2D(m, border) of an H x W image (rows first, as
// ToMatrix builds it); out is (H + 2 border) x (W + 2 border).  Returns 1
// when the reference rejects the arguments (it throws for a border wider than
// the image).
int ref_mirror(const double* in, int W, int H, int border, double* out) try {
  const unsigned long h = unsigned(H), w = unsigned(W);
  matrix<> m(h, w);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) m(unsigned(y), unsigned(x)) = in[size_t(y) * W + x];
  matrix<> o = lib_image::border_mirror_2D(m, unsigned(border));
  const int Wo = W + 2 * border, Ho = H + 2 * border;
  for (int y = 0; y < Ho; y++)
    for (int x = 0; x < Wo; x++) out[size_t(y) * Wo + x] = o(unsigned(y), unsigned(x));
  return 0;
} catch (...) {
  return 1;
}

// lib_image::hist_gradient_2D(labels, r, 8, kernel) with the default
// chi-squared distance (mex_pb_parts:160-227); kernel = gaussian(ksigma), or
// none when ksigma <= 0 (the texton call).  labels H x W; out[8][H][W].
void ref_hist_gradient(const unsigned char* labels, int W, int H, int r, double ksigma, double* out) {
  const unsigned long h = unsigned(H), w = unsigned(W);
  matrix<unsigned long> L(h, w);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) L(unsigned(y), unsigned(x)) = labels[size_t(y) * W + x];
  auto_collection<matrix<>, array_list<matrix<> > > g =
      ksigma > 0 ? lib_image::hist_gradient_2D(L, unsigned(r), 8, lib_image::gaussian(ksigma))
                 : lib_image::hist_gradient_2D(L, unsigned(r), 8);
  for (int o = 0; o < 8; o++)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) out[(size_t(o) * H + y) * W + x] = (*g)[unsigned(o)](unsigned(y), unsigned(x));
}

// nonmax_channels (max and argmax over the 8 orientations, angle i*PI/8 with
// PI = 3.1415926535 from include/UCM_CONSTANTS.h:4, the running maximum
// restarting at the smallest positive double and the argmax carried over from
// the previous pixel, as written) -> lib_image::nonmax_oriented_2D(a, ind,
// PI/8) -> multiscalePb's clip.  mpb_all[8][H][W] -> nmax[H][W].
void ref_nonmax(const double* mpb_all, int W, int H, double* nmax) {
  const double PI = 3.1415926535;
  const size_t N = size_t(W) * H;
  const unsigned long h = unsigned(H), w = unsigned(W);
  matrix<> a(h, w), ind(h, w);
  double maxVal = std::numeric_limits<double>::min();
  int maxValIndex = 0;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      for (int i = 0; i < 8; i++) {
        const double d = mpb_all[size_t(i) * N + size_t(y) * W + x];
        if (d > maxVal) {
          maxVal = d;
          maxValIndex = i;
        }
      }
      if (maxVal < 0) maxVal = 0;
      a(unsigned(y), unsigned(x)) = maxVal;
      ind(unsigned(y), unsigned(x)) = double(maxValIndex) * PI / 8.0;
      maxVal = std::numeric_limits<double>::min();
    }
  matrix<> nm = lib_image::nonmax_oriented_2D(a, ind, PI / 8);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      double v = nm(unsigned(y), unsigned(x));
      if (1.2 * v > 1) v = 1;
      if (v < 0) v = 0;
      nmax[size_t(y) * W + x] = v;
    }
}

}  // extern "C"

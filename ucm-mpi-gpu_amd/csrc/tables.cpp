// Host construction of the constant tables the gPb kernels read.  Each table
// restates the reference routine it follows (file:line cited at each builder);
// everything here is computed once per process and cached.
#include <algorithm>
#include <cassert>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>

#include "gpb_common.hpp"

namespace gpb {

const double kMPbWeights[kNumCues] = {0.0146, 0.0145, 0.0163, 0.0210, 0.0243, 0.0287,
                                      0.0166, 0.0185, 0.0204, 0.0101, 0.0111, 0.0141};
const double kGPbWeights[kNumCues + 1] = {0,      0,      0.0039, 0.0050, 0.0058, 0.0069, 0.0040,
                                          0.0044, 0.0049, 0.0024, 0.0027, 0.0170, 0.0074};
const int kRadiusBg[3] = {3, 5, 10};
const int kRadiusCg[3] = {5, 10, 20};

// ---------------------------------------------------------------------------
// Schmid bank: SPFeatures::makefilter / makeSfilters (SPFeatures.cpp:117-183).
// F = cos(r * (pi*tau/sigma)) * exp(-r^2 / (2 sigma^2)); subtract the mean;
// divide by the L1 norm taken before the mean removal (a quirk kept as written).
static void schmid_kernel(int sigma, int tau, double* out /* 49*49 */) {
  const int sup = kSchmidSup, hsup = kSchmidHalf;
  double fsum = 0, fsum_abs = 0;
  for (int i = 0; i < sup; i++) {
    for (int j = 0; j < sup; j++) {
      const double x = j - hsup, y = i - hsup;
      const double r = std::sqrt(x * x + y * y);
      double f = std::cos(r * (kRefPI * tau / sigma));
      f *= std::exp(-r * r / (2 * sigma * sigma));
      out[i * sup + j] = f;
      fsum += f;
      fsum_abs += std::fabs(f);
    }
  }
  const double mean = fsum / (sup * sup);
  for (int n = 0; n < sup * sup; n++) out[n] = (out[n] - mean) / fsum_abs;
}

static SchmidTable build_schmid() {
  static const int sig_tau[kSchmidNF][2] = {{2, 1}, {4, 1}, {4, 2}, {6, 1},  {6, 2},  {6, 3}, {8, 1},
                                            {8, 2}, {8, 3}, {10, 1}, {10, 2}, {10, 3}, {10, 4}};
  SchmidTable t;
  const int sup = kSchmidSup, hsup = kSchmidHalf;
  t.bank.resize(size_t(kSchmidNF) * sup * sup);
  for (int f = 0; f < kSchmidNF; f++) schmid_kernel(sig_tau[f][0], sig_tau[f][1], &t.bank[size_t(f) * sup * sup]);
  std::map<int, std::vector<std::pair<int, int>>> cls;
  for (int dy = -hsup; dy <= hsup; dy++)
    for (int dx = -hsup; dx <= hsup; dx++) cls[dx * dx + dy * dy].push_back({dx, dy});
  t.num_classes = int(cls.size());
  t.class_start.push_back(0);
  t.weight.assign(size_t(kSchmidNF) * t.num_classes, 0.f);
  int c = 0;
  for (auto& kv : cls) {
    for (auto& m : kv.second) {
      t.member_dx.push_back(m.first);
      t.member_dy.push_back(m.second);
    }
    t.class_start.push_back(int(t.member_dx.size()));
    const auto& m0 = kv.second.front();
    for (int f = 0; f < kSchmidNF; f++) {
      const float w = float(t.bank[size_t(f) * sup * sup + (m0.second + hsup) * sup + (m0.first + hsup)]);
      for (auto& m : kv.second) {
        const float wm = float(t.bank[size_t(f) * sup * sup + (m.second + hsup) * sup + (m.first + hsup)]);
        if (wm != w) throw std::logic_error("Schmid kernel is not radially symmetric");
      }
      t.weight[size_t(f) * t.num_classes + c] = w;
    }
    c++;
  }
  return t;
}

const SchmidTable& schmid_table() {
  static const SchmidTable t = build_schmid();
  return t;
}

void schmid_response_at(const uint32_t* packed, int width, int height, int x, int y, float* r13) {
  const SchmidTable& t = schmid_table();
  for (int f = 0; f < kSchmidNF; f++) r13[f] = 0.f;
  for (int c = 0; c < t.num_classes; c++) {
    float s = 0.f;
    for (int m = t.class_start[c]; m < t.class_start[c + 1]; m++) {
      const int gx = reflect_index(x + t.member_dx[m], width);
      const int gy = reflect_index(y + t.member_dy[m], height);
      s += float(grey_u8(packed[size_t(gy) * width + gx]));
    }
    for (int f = 0; f < kSchmidNF; f++) r13[f] = std::fmaf(t.weight[size_t(f) * t.num_classes + c], s, r13[f]);
  }
}

// ---------------------------------------------------------------------------
// Orientation slices: orientation_slice_map (lib_image.cc:3052-3079) evaluates
// atan2 at window coordinates shifted by half a pixel, and
// compute_hist_gradient_2D (lib_image.cc:3120-3231) pairs window position o
// with the label at centre + (r - o), so a label at relative (drow, dcol)
// falls in the slice computed at window position (r - drow, r - dcol).
int slice_of_offset(int drow, int dcol) {
  const double xr = double(-drow) - 0.5;  // window x: n_x - size/2 with n_x = r - drow
  const double yc = double(-dcol) - 0.5;
  const long double ori = std::atan2((double)yc, (double)xr) + M_PIl;
  long idx = (long)std::floor((double)(ori / M_PIl * (long double)kNumOri));
  if (idx >= 2 * kNumOri) idx = 2 * kNumOri - 1;
  return int(idx);
}

// lib_image::gaussian(sigma, deriv 0, no Hilbert): support ceil(3 sigma), unit L1.
std::vector<double> gaussian_1d(double sigma) {
  const long support = (long)std::ceil(3 * sigma);
  const double sigma2_inv = 1.0 / (sigma * sigma);
  const double neg_two_sigma2_inv = -0.5 * sigma2_inv;
  std::vector<double> m(2 * support + 1);
  double x = -double(support);
  for (size_t n = 0; n < m.size(); n++, x++) m[n] = std::exp(x * x * neg_two_sigma2_inv);
  double s = 0;
  for (double v : m) s += std::fabs(v);
  for (double& v : m) v /= s;
  return m;
}

// ---------------------------------------------------------------------------
// Oriented nonmax: nonmax_oriented_2D (lib_image.cc:4083-4200) with the
// per-pixel orientation k*PI/8 from nonmax_channels.cpp:14 (truncated PI) and
// tolerance PI/8.  Branch codes: 0 theta==0, 1 theta<pi/4, 2 theta<pi/2,
// 3 theta==pi/2, 4 theta<3pi/4, 5 otherwise.
static NonmaxTable build_nonmax() {
  NonmaxTable t{};
  const double tol = kRefPI / 8;
  double oris[kNumOri];
  for (int k = 0; k < kNumOri; k++) oris[k] = double(k) * kRefPI / double(kNumOri);
  for (int k = 0; k < kNumOri; k++) {
    const double ori = oris[k];
    double theta = (double)((long double)ori + M_PI_2l);
    theta = (double)((long double)theta - std::floor((long double)theta / M_PIl) * M_PIl);
    double d = 0;
    int br;
    if (theta == 0) {
      br = 0;
    } else if ((long double)theta < M_PI_4l) {
      br = 1;
      d = (double)std::tan((long double)theta);
    } else if ((long double)theta < M_PI_2l) {
      br = 2;
      d = (double)std::tan(M_PI_2l - (long double)theta);
    } else if ((long double)theta == M_PI_2l) {
      br = 3;
    } else if ((long double)theta < 3.0L * M_PI_4l) {
      br = 4;
      d = (double)std::tan((long double)theta - M_PI_2l);
    } else {
      br = 5;
      d = (double)std::tan(M_PIl - (long double)theta);
    }
    t.branch[k] = br;
    t.d[k] = d;
    for (int j = 0; j < kNumOri; j++) {
      double od = oris[j] - ori;
      od = (double)((long double)od - std::floor((long double)od / (2 * M_PIl)) * (2 * M_PIl));
      if ((long double)od >= M_PIl) od = (double)(2 * M_PIl - (long double)od);
      if ((long double)od >= M_PI_2l) od = (double)(M_PIl - (long double)od);
      od = (od <= tol) ? 0 : (od - tol);
      t.cosdiff[k][j] = std::cos(od);
    }
  }
  return t;
}

const NonmaxTable& nonmax_table() {
  static const NonmaxTable t = build_nonmax();
  return t;
}

// ---------------------------------------------------------------------------
// sPb filters.  The reference CPU oeFilter (oeFilter.cpp:21-196) reads
// uninitialised memory, so the filters follow the GPU construction instead:
// gaussian_2D (damascene/src/textons/filters.cpp:264-341) with sigma 1,
// support 3 and first derivative: a separable Gaussian x Gaussian-derivative
// kernel on a support grid enlarged for rotation, rotated by bilinear
// interpolation (compute_rotate_2D, filters.cpp:209-255) onto a 7x7 grid, then
// zero mean and unit L1 norm.  Rows are the slow image axis (lib_image's x).
// Orientation o uses rotation o*pi/8, which differentiates along the normal of
// an edge running at angle o*pi/8 in the (row, col) plane, the same angle at
// which hist_gradient_2D orientation o splits its disc.
static std::vector<double> gauss1d_support(double sigma, int deriv, long support) {
  const double sigma2_inv = 1.0 / (sigma * sigma);
  const double neg_two_sigma2_inv = -0.5 * sigma2_inv;
  const long size = 2 * support + 1;
  std::vector<double> m(size);
  double x = -double(support);
  for (long n = 0; n < size; n++, x++) {
    m[n] = std::exp(x * x * neg_two_sigma2_inv);
    if (deriv == 1) m[n] *= -x;
  }
  if (deriv > 0) {
    double s = 0;
    for (double v : m) s += v;
    const double mean = s / double(size);
    for (double& v : m) v -= mean;
  }
  double sabs = 0;
  for (double v : m) sabs += std::fabs(v);
  for (double& v : m) v /= sabs;
  return m;
}

static long rotated_support(double sa, double sb, double ori, bool x_axis) {
  double a0, a1;
  if (x_axis) {
    const double c = sa * std::cos(ori), s = sb * std::sin(ori);
    a0 = std::fabs(c - s);
    a1 = std::fabs(c + s);
  } else {
    const double s = sa * std::sin(ori), c = sb * std::cos(ori);
    a0 = std::fabs(s - c);
    a1 = std::fabs(s + c);
  }
  return (long)std::ceil(a0 > a1 ? a0 : a1);
}

static std::vector<double> build_spb_filters() {
  std::vector<double> out(size_t(kNumOri) * kSPbSize * kSPbSize, 0.0);
  for (int o = 0; o < kNumOri; o++) {
    const double ori = double(o) * M_PI / kNumOri;
    const long sxr = rotated_support(kSPbRadius, kSPbRadius, -ori, true);
    const long syr = rotated_support(kSPbRadius, kSPbRadius, -ori, false);
    const std::vector<double> mx = gauss1d_support(1.0, 0, sxr);
    const std::vector<double> my = gauss1d_support(1.0, 1, syr);
    const long nx = 2 * sxr + 1, ny = 2 * syr + 1;
    std::vector<double> m(size_t(nx * ny));
    for (long i = 0; i < nx; i++)
      for (long j = 0; j < ny; j++) m[size_t(i * ny + j)] = mx[size_t(i)] * my[size_t(j)];
    double* dst = &out[size_t(o) * kSPbSize * kSPbSize];
    const double co = std::cos(ori), so = std::sin(ori);
    const double ox = double(nx - 1) / 2, oy = double(ny - 1) / 2;
    for (int a = 0; a < kSPbSize; a++) {
      const double u = double(a) - double(kSPbSize - 1) / 2;
      for (int b = 0; b < kSPbSize; b++) {
        const double v = double(b) - double(kSPbSize - 1) / 2;
        const double x = u * co + v * so + ox;
        const double y = v * co - u * so + oy;
        double val = 0;
        if (x >= 0 && y >= 0) {
          const long x0 = (long)std::floor(x), x1 = (long)std::ceil(x);
          const long y0 = (long)std::floor(y), y1 = (long)std::ceil(y);
          if (x1 < nx && y1 < ny) {
            const double m00 = m[size_t(x0 * ny + y0)], m01 = m[size_t(x0 * ny + y1)];
            const double m10 = m[size_t(x1 * ny + y0)], m11 = m[size_t(x1 * ny + y1)];
            const double t0 = (x0 != x1) ? ((x1 - x) * m00 + (x - x0) * m10) : m00;
            const double t1 = (x0 != x1) ? ((x1 - x) * m01 + (x - x0) * m11) : m01;
            val = (y0 != y1) ? ((y1 - y) * t0 + (y - y0) * t1) : t0;
          }
        }
        dst[a * kSPbSize + b] = val;
      }
    }
    double s = 0;
    for (int n = 0; n < kSPbSize * kSPbSize; n++) s += dst[n];
    const double mean = s / (kSPbSize * kSPbSize);
    double sabs = 0;
    for (int n = 0; n < kSPbSize * kSPbSize; n++) {
      dst[n] -= mean;
      sabs += std::fabs(dst[n]);
    }
    for (int n = 0; n < kSPbSize * kSPbSize; n++) dst[n] /= sabs;
  }
  return out;
}

const std::vector<double>& spb_filters() {
  static const std::vector<double> f = build_spb_filters();
  return f;
}

// ---------------------------------------------------------------------------
// Intervening-contour programs.  These restate Group::interveningContour,
// Group::ic_walk (buildW/src/ic.cc:86-366, 375-560) and the support lookup of
// Group::computeAffinities2 (buildW/src/affinity.cc:58-156) symbolically: the
// walk is replayed on a virtual image whose clipping matches the pixel's, and
// instead of lattice values each recorded point keeps the list of pixels whose
// mPb the walk maximised over.  Lattice H(x,y) is mPb(row y-1, col x) and
// V(x,y) is mPb(row y, col x-1) (spectralPb.cpp:55-59, buildW.cpp:57-117).
namespace {

struct VPoint {
  int x, y;
  std::vector<std::pair<int, int>> sites;  // (col, row) absolute in the virtual image
};

struct Walker {
  int width, height, x0, y0, wr;
  std::vector<std::vector<VPoint>> rows;  // scan lines, index y - y0 + wr

  void add_site_H(std::vector<std::pair<int, int>>& s, int x, int y) { s.push_back({x, y - 1}); }
  void add_site_V(std::vector<std::pair<int, int>>& s, int x, int y) { s.push_back({x - 1, y}); }

  void walk(int x1, int y1, int x2, int y2, int x3, int y3) {
    const long long dx1 = x1 - x0, dy1 = y1 - y0, dx2 = x2 - x0, dy2 = y2 - y0, dx3 = x3 - x0, dy3 = y3 - y0;
    const long long dot11 = dx1 * dx1 + dy1 * dy1, dot22 = dx2 * dx2 + dy2 * dy2, dot33 = dx3 * dx3 + dy3 * dy3;
    const int dx = x2 - x0, dy = y2 - y0, adx = std::abs(dx), ady = std::abs(dy);
    // Major axis per quadrant, exactly as the octant table of ic_walk.
    bool major_x;
    bool forward;  // octants 0,1,6,7 keep walk order; 2..5 are appended reversed
    if (dx > 0 && dy >= 0) {
      major_x = adx > ady;
      forward = true;
    } else if (dx <= 0 && dy > 0) {
      major_x = !(adx < ady);
      forward = false;
    } else if (dy <= 0 && dx < 0) {
      major_x = adx > ady;
      forward = false;
    } else if (dx >= 0 && dy < 0) {
      major_x = !(adx < ady);
      forward = true;
    } else {
      throw std::logic_error("ic walk: degenerate direction");
    }
    const int sx = dx > 0 ? 1 : (dx < 0 ? -1 : 0);
    const int sy = dy > 0 ? 1 : (dy < 0 ? -1 : 0);
    int t = major_x ? -adx : -ady;
    int xi = x0, yi = y0, oldx = x0, oldy = y0;
    std::vector<std::pair<int, int>> acc;  // sites seen so far (running max set)
    std::vector<VPoint> pts;
    while (xi != x2 || yi != y2) {
      if (major_x) {
        xi += sx;
        t += ady << 1;
        if (t > 0) {
          yi += sy;
          t -= adx << 1;
        }
      } else {
        yi += sy;
        t += adx << 1;
        if (t > 0) {
          xi += sx;
          t -= ady << 1;
        }
      }
      const long long dxi = xi - x0, dyi = yi - y0;
      const long long dotii = dxi * dxi + dyi * dyi;
      const long long doti1 = dxi * dx1 + dyi * dy1;
      const long long doti2 = dxi * dx2 + dyi * dy2;
      const long long doti3 = dxi * dx3 + dyi * dy3;
      const bool good = (doti2 * doti2 == dotii * dot22) ||
                        (dot11 * doti2 * doti2 > dot22 * doti1 * doti1 && dot33 * doti2 * doti2 >= dot22 * doti3 * doti3);
      if (oldx == xi) {
        if (yi > oldy) add_site_H(acc, xi, yi);
        else if (yi < oldy) add_site_H(acc, oldx, oldy);
      } else if (oldy == yi) {
        if (xi > oldx) add_site_V(acc, xi, yi);
        else if (xi < oldx) add_site_V(acc, oldx, oldy);
      } else if (xi > oldx && yi > oldy) {
        add_site_H(acc, oldx, yi); add_site_H(acc, xi, yi); add_site_V(acc, xi, oldy); add_site_V(acc, xi, yi);
      } else if (xi > oldx && yi < oldy) {
        add_site_H(acc, oldx, oldy); add_site_H(acc, xi, oldy); add_site_V(acc, xi, oldy); add_site_V(acc, xi, yi);
      } else if (xi < oldx && yi > oldy) {
        add_site_H(acc, oldx, yi); add_site_H(acc, xi, yi); add_site_V(acc, oldx, oldy); add_site_V(acc, oldx, yi);
      } else {
        add_site_H(acc, oldx, oldy); add_site_H(acc, xi, oldy); add_site_V(acc, oldx, oldy); add_site_V(acc, oldx, yi);
      }
      oldx = xi;
      oldy = yi;
      if (!good) continue;
      VPoint p{xi, yi, acc};
      pts.push_back(std::move(p));
    }
    if (forward) {
      for (auto& p : pts) rows[size_t(p.y - y0 + wr)].push_back(p);
    } else {
      for (auto it = pts.rbegin(); it != pts.rend(); ++it) rows[size_t(it->y - y0 + wr)].push_back(*it);
    }
  }

  void corner(int x, int y) {
    VPoint p{x, y, {}};
    rows[size_t(y - y0 + wr)].push_back(p);
  }

  // Order of walks as in Group::interveningContour (ic.cc:400-511).
  void run() {
    rows.assign(size_t(2 * wr + 1), {});
    const int rxa = std::max(0, x0 - wr), rya = std::max(0, y0 - wr);
    const int rxb = std::min(x0 + wr, width - 1), ryb = std::min(y0 + wr, height - 1);
    if (x0 > rxa) {
      if (y0 > 0 && y0 < ryb) walk(rxa, y0 - 1, rxa, y0, rxa, y0 + 1);
      for (int y = y0 - 1; y > rya; y--) walk(rxa, y - 1, rxa, y, rxa, y + 1);
    }
    if (x0 > rxa + 1 || y0 > rya + 1 || (x0 > rxa && y0 > rya)) walk(rxa, rya + 1, rxa, rya, rxa + 1, rya);
    if ((x0 == rxa && y0 == rya + 1) || (x0 == rxa + 1 && y0 == rya)) corner(rxa, rya);
    if (y0 > rya)
      for (int x = rxa + 1; x < rxb; x++) walk(x - 1, rya, x, rya, x + 1, rya);
    if (y0 > rya + 1 || x0 < rxb - 1 || (y0 > rya && x0 < rxb)) walk(rxb - 1, rya, rxb, rya, rxb, rya + 1);
    if ((x0 == rxb - 1 && y0 == rya) || (x0 == rxb && y0 == rya + 1)) corner(rxb, rya);
    if (x0 < rxb)
      for (int y = rya + 1; y < y0; y++) walk(rxb, y - 1, rxb, y, rxb, y + 1);
    if (x0 > rxa)
      for (int y = y0 + 1; y < ryb; y++) walk(rxa, y - 1, rxa, y, rxa, y + 1);
    if (x0 > rxa + 1 || y0 < ryb - 1 || (x0 > rxa && y0 < ryb)) walk(rxa, ryb - 1, rxa, ryb, rxa + 1, ryb);
    if ((x0 == rxa && y0 == ryb - 1) || (x0 == rxa + 1 && y0 == ryb)) corner(rxa, ryb);
    if (y0 < ryb)
      for (int x = rxa + 1; x < rxb; x++) walk(x - 1, ryb, x, ryb, x + 1, ryb);
    if (y0 < ryb - 1 || x0 < rxb - 1 || (y0 < ryb && x0 < rxb)) walk(rxb - 1, ryb, rxb, ryb, rxb, ryb - 1);
    if ((x0 == rxb - 1 && y0 == ryb) || (x0 == rxb && y0 == ryb - 1)) corner(rxb, ryb);
    if (x0 < rxb) {
      for (int y = ryb - 1; y > y0; y--) walk(rxb, y - 1, rxb, y, rxb, y + 1);
      if (y0 > 0 && y0 < ryb) walk(rxb, y0 - 1, rxb, y0, rxb, y0 + 1);
    }
    for (auto& r : rows)
      for (size_t i = 1; i < r.size(); i++)
        if (!(r[i - 1].x < r[i].x)) throw std::logic_error("ic walk: scan line not strictly increasing");
  }
};

ICProgram build_program(int left, int right, int top, int bottom) {
  const int wr = kICRadius;
  Walker w;
  w.wr = wr;
  w.x0 = left;
  w.y0 = top;
  w.width = left + right + 1;
  w.height = top + bottom + 1;
  // An unclipped side needs at least one more column/row beyond the window so
  // that min(x0 + wr, width - 1) == x0 + wr; a window edge exactly on the image
  // edge is indistinguishable, so this choice is exact.
  w.run();
  std::vector<VPoint> support;
  for (auto& r : w.rows)
    for (auto& p : r) support.push_back(p);
  ICProgram prog;
  // Support lookup of computeAffinities2, including its cursor that may skip
  // entries of the next scan line.
  size_t idx = 0;
  const int x = w.x0, y = w.y0;
  for (int u = -wr; u <= wr; u++) {
    const int yy = y + u;
    for (int v = -wr; v <= wr; v++) {
      const int xx = x + v;
      if (xx < 0 || xx >= w.width) continue;
      if (yy < 0 || yy >= w.height) continue;
      if (u * u + v * v > wr * wr) continue;
      while (idx < support.size() && support[idx].y < yy) idx++;
      while (idx < support.size() && support[idx].x < xx) idx++;
      if (u == 0 && v == 0) continue;
      ICConnection c{};
      c.dx = int8_t(v);
      c.dy = int8_t(u);
      c.site_start = uint16_t(prog.site_dx.size());
      c.site_count = 0;
      c.found = 0;
      if (idx < support.size() && support[idx].x == xx && support[idx].y == yy) {
        c.found = 1;
        std::vector<std::pair<int, int>> s = support[idx].sites;
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        for (auto& st : s) {
          prog.site_dx.push_back(int8_t(st.first - x));
          prog.site_dy.push_back(int8_t(st.second - y));
        }
        c.site_count = uint8_t(s.size());
        idx++;
      }
      prog.conn.push_back(c);
    }
  }
  return prog;
}

}  // namespace

const ICProgram& ic_program(int left, int right, int top, int bottom) {
  static std::mutex mu;
  static std::map<int, ICProgram> cache;
  const int key = ((left * 6 + right) * 6 + top) * 6 + bottom;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  auto res = cache.emplace(key, build_program(left, right, top, bottom));
  return res.first->second;
}

void disc_offsets(int* dx, int* dy) {
  int n = 0;
  for (int u = -kICRadius; u <= kICRadius; u++)
    for (int v = -kICRadius; v <= kICRadius; v++)
      if (u * u + v * v <= kICRadius * kICRadius) {
        dx[n] = v;
        dy[n] = u;
        n++;
      }
}

void upper_offsets(int* dx, int* dy) {
  int n = 0;
  for (int u = 0; u <= kICRadius; u++)
    for (int v = -kICRadius; v <= kICRadius; v++)
      if (u * u + v * v <= kICRadius * kICRadius && (u > 0 || v > 0)) {
        dx[n] = v;
        dy[n] = u;
        n++;
      }
}

}  // namespace gpb

// Dense symmetric eigensolver (FP64) for the small projected matrices of the
// Lanczos solver: Householder reduction to tridiagonal form followed by the
// implicit-shift QL iteration (Wilkinson & Reinsch, Handbook for Automatic
// Computation vol. II, procedures tred2 and tql2).  Replaces the ACML
// dstebz/dstein calls of the reference GPU solver (lanczos.cu:297-380).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace gpb {

// a: n x n row-major symmetric (destroyed).  On return w holds the eigenvalues
// in descending order and z (row-major n x n) the matching eigenvectors as
// columns.
void symeig_desc(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& z) {
  std::vector<double> d(n), e(n);
  z = a;  // symmetric: the same array read column-major
  // Z is held column-major: the inner loops below run down columns (the
  // Householder updates and the QL rotations of two columns), which are then
  // contiguous.
  auto Z = [&](int i, int j) -> double& { return z[size_t(j) * n + i]; };
  // Householder tridiagonalisation (accumulating the transformation in Z)
  for (int j = 0; j < n; j++) d[j] = Z(n - 1, j);
  for (int i = n - 1; i > 0; i--) {
    double scale = 0.0, h = 0.0;
    for (int k = 0; k < i; k++) scale += std::fabs(d[k]);
    if (scale == 0.0) {
      e[i] = d[i - 1];
      for (int j = 0; j < i; j++) {
        d[j] = Z(i - 1, j);
        Z(i, j) = 0.0;
        Z(j, i) = 0.0;
      }
    } else {
      for (int k = 0; k < i; k++) {
        d[k] /= scale;
        h += d[k] * d[k];
      }
      double f = d[i - 1];
      double g = std::sqrt(h);
      if (f > 0) g = -g;
      e[i] = scale * g;
      h -= f * g;
      d[i - 1] = f - g;
      for (int j = 0; j < i; j++) e[j] = 0.0;
      for (int j = 0; j < i; j++) {
        f = d[j];
        Z(j, i) = f;
        g = e[j] + Z(j, j) * f;
        for (int k = j + 1; k <= i - 1; k++) {
          g += Z(k, j) * d[k];
          e[k] += Z(k, j) * f;
        }
        e[j] = g;
      }
      f = 0.0;
      for (int j = 0; j < i; j++) {
        e[j] /= h;
        f += e[j] * d[j];
      }
      const double hh = f / (h + h);
      for (int j = 0; j < i; j++) e[j] -= hh * d[j];
      for (int j = 0; j < i; j++) {
        f = d[j];
        g = e[j];
        for (int k = j; k <= i - 1; k++) Z(k, j) -= (f * e[k] + g * d[k]);
        d[j] = Z(i - 1, j);
        Z(i, j) = 0.0;
      }
    }
    d[i] = h;
  }
  for (int i = 0; i < n - 1; i++) {
    Z(n - 1, i) = Z(i, i);
    Z(i, i) = 1.0;
    const double h = d[i + 1];
    if (h != 0.0) {
      for (int k = 0; k <= i; k++) d[k] = Z(k, i + 1) / h;
      for (int j = 0; j <= i; j++) {
        double g = 0.0;
        for (int k = 0; k <= i; k++) g += Z(k, i + 1) * Z(k, j);
        for (int k = 0; k <= i; k++) Z(k, j) -= g * d[k];
      }
    }
    for (int k = 0; k <= i; k++) Z(k, i + 1) = 0.0;
  }
  for (int j = 0; j < n; j++) {
    d[j] = Z(n - 1, j);
    Z(n - 1, j) = 0.0;
  }
  Z(n - 1, n - 1) = 1.0;
  e[0] = 0.0;
  // implicit QL on the tridiagonal (d, e)
  for (int i = 1; i < n; i++) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  double f = 0.0, tst1 = 0.0;
  const double eps = std::ldexp(1.0, -52);
  for (int l = 0; l < n; l++) {
    tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
    int m = l;
    while (m < n) {
      if (std::fabs(e[m]) <= eps * tst1) break;
      m++;
    }
    if (m == n) m = n - 1;
    if (m > l) {
      int iter = 0;
      do {
        if (++iter > 200) throw std::runtime_error("symeig: QL iteration did not converge");
        double g = d[l];
        double p = (d[l + 1] - g) / (2.0 * e[l]);
        double r = std::hypot(p, 1.0);
        if (p < 0) r = -r;
        d[l] = e[l] / (p + r);
        d[l + 1] = e[l] * (p + r);
        const double dl1 = d[l + 1];
        double h = g - d[l];
        for (int i = l + 2; i < n; i++) d[i] -= h;
        f += h;
        p = d[m];
        double c = 1.0, c2 = c, c3 = c, el1 = e[l + 1], s = 0.0, s2 = 0.0;
        for (int i = m - 1; i >= l; i--) {
          c3 = c2;
          c2 = c;
          s2 = s;
          g = c * e[i];
          h = c * p;
          r = std::hypot(p, e[i]);
          e[i + 1] = s * r;
          s = e[i] / r;
          c = p / r;
          p = c * d[i] - s * g;
          d[i + 1] = h + s * (c * g + s * d[i]);
          for (int k = 0; k < n; k++) {
            h = Z(k, i + 1);
            Z(k, i + 1) = s * Z(k, i) + c * h;
            Z(k, i) = c * Z(k, i) - s * h;
          }
        }
        p = -s * s2 * c3 * el1 * e[l] / dl1;
        e[l] = s * p;
        d[l] = c * p;
      } while (std::fabs(e[l]) > eps * tst1);
    }
    d[l] += f;
    e[l] = 0.0;
  }
  // sort descending
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a_, int b_) { return d[a_] > d[b_]; });
  w.resize(n);
  std::vector<double> zs(size_t(n) * n);
  for (int j = 0; j < n; j++) {
    w[j] = d[idx[j]];
    for (int k = 0; k < n; k++) zs[size_t(k) * n + j] = Z(k, idx[j]);
  }
  z.swap(zs);
}

}  // namespace gpb

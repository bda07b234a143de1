// Test-infrastructure driver around the reference's own buildW kernels
// (buildW/src/ic.cc, affinity.cc, smatrix.cc, compiled from the reference
// tree by `make ref`): builds the dual lattice from an mPb map exactly as
// src/spectralPb.cpp:55-59 and src/buildW.cpp:70-153 do, runs
// Group::computeSupport + Group::computeAffinities2 (with the reference's
// symmetrize), and returns the whole sparse W row by row with its row sums.
// Used only by tests to pin the oracle's intervening-contour affinity.
#include "affinity.hh"
#include "ic.hh"
#include "smatrix.hh"

extern "C" {

// mpb: H x W row-major.  On entry *nnz is the capacity of rows/cols/vals;
// on return, the number of entries (or the needed capacity, returning 1).
int ref_build_w(int W, int H, const double* mpb, int* nnz, int* rows, int* cols, double* vals, double* rowsum) {
  Group::DualLattice b;
  // l1 = (H+1) x W with mPb in rows 1..H; H(x, y) = l1(y, x)  (buildW.cpp:70-110)
  b.H.resize(W, H + 1);
  for (int x = 0; x < W; x++)
    for (int y = 0; y <= H; y++) b.H(x, y) = y == 0 ? 0.0f : float(mpb[size_t(y - 1) * W + x]);
  // l2 = H x (W+1) with mPb in columns 1..W; V(x, y) = l2(y, x)  (buildW.cpp:112-149)
  b.V.resize(W + 1, H);
  for (int x = 0; x <= W; x++)
    for (int y = 0; y < H; y++) b.V(x, y) = x == 0 ? 0.0f : float(mpb[size_t(y) * W + (x - 1)]);
  b.width = b.H.size(0);
  b.height = b.V.size(1);
  Group::SupportMap ic;
  Group::computeSupport(b, 5, 1.0f, ic);  // buildW.cpp:24-25,151
  SMatrix* S = nullptr;
  Group::computeAffinities2(ic, 0.1f, 5, &S);  // buildW.cpp:156
  if (!S) return 2;
  int total = 0;
  for (int r = 0; r < S->n; r++) total += S->nz[r];
  if (total > *nnz) {
    *nnz = total;
    delete S;
    return 1;
  }
  int k = 0;
  for (int r = 0; r < S->n; r++) {
    double s = 0;
    for (int i = 0; i < S->nz[r]; i++) {
      rows[k] = r;
      cols[k] = S->col[r][i];
      vals[k] = S->values[r][i];
      s += S->values[r][i];
      k++;
    }
    rowsum[r] = s;
  }
  *nnz = total;
  delete S;
  return 0;
}

}  // extern "C"

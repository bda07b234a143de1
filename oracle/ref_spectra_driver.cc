// Driver around the reference's CPU eigensolver, used as the CPU baseline's
// eigensolver and nothing else: Spectra's SymGEigsSolver (Spectra/, header
// only, with the Eigen 3.2.8 the reference vendors under Eigen/; compiled from
// the reference tree by `make ref`) set up the way src/spectralPb.cpp:196-244
// sets it up -- A = D - W as an Eigen sparse matrix behind SparseGenMatProd,
// B = D behind SparseCholesky, SMALLEST_MAGN selection, GEIGS_CHOLESKY mode,
// init() with its fixed seed, compute(maxit, tol).  One deviation, declared:
// A holds the full symmetric D - W, where the reference exports only the
// entries with column >= row (src/buildW.cpp:199-208) and so multiplies a
// triangle as a general matrix.
#include <algorithm>
#include <vector>

#include <Eigen/Sparse>

#include "Spectra/MatOp/SparseCholesky.h"
#include "Spectra/MatOp/SparseGenMatProd.h"
#include "Spectra/SymGEigsSolver.h"

extern "C" {

// W: nnz triplets (rows, cols, vals) of the symmetric affinity including its
// diagonal; deg[n] its row sums (D).  Returns the number of converged pairs
// (Spectra reports only those); evals[0..nconv) ascending and evecs[nconv][n]
// (B-normalised, v' D v = 1) in the same order; iters / ops: Spectra's
// restart and operator counts.  Spectra's threshold tol * max(eps^(2/3),
// |theta|) asks the trivial theta = 0 pair for a residual near 1e-21, which
// it meets only when the constant vector is captured exactly; otherwise
// compute() runs all maxit restarts and reports fewer than nev pairs, as the
// reference's call does (src/spectralPb.cpp:236-244 then logs "Geigs is
// failing").
int ref_spectra_geigs(int n, int nnz, const int* rows, const int* cols, const double* vals, const double* deg,
                      int nev, int ncv, int maxit, double tol, double* evals, double* evecs, int* iters, int* ops) {
  typedef Eigen::Triplet<double> T;
  std::vector<T> a;
  a.reserve(size_t(nnz) + size_t(n));
  for (int i = 0; i < n; i++) a.push_back(T(i, i, deg[i]));
  for (int k = 0; k < nnz; k++) a.push_back(T(rows[k], cols[k], -vals[k]));
  Eigen::SparseMatrix<double> A(n, n), B(n, n);
  A.setFromTriplets(a.begin(), a.end());
  A.makeCompressed();
  std::vector<T> b;
  b.reserve(size_t(n));
  for (int i = 0; i < n; i++) b.push_back(T(i, i, deg[i]));
  B.setFromTriplets(b.begin(), b.end());
  B.makeCompressed();
  Spectra::SparseGenMatProd<double> opA(A);
  Spectra::SparseCholesky<double> opB(B);
  Spectra::SymGEigsSolver<double, Spectra::SMALLEST_MAGN, Spectra::SparseGenMatProd<double>,
                          Spectra::SparseCholesky<double>, Spectra::GEIGS_CHOLESKY>
      geigs(&opA, &opB, nev, ncv);
  geigs.init();
  const int nconv = geigs.compute(maxit, tol);
  *iters = geigs.num_iterations();
  *ops = geigs.num_operations();
  Eigen::VectorXd ev = geigs.eigenvalues();
  Eigen::MatrixXd vec = geigs.eigenvectors();
  std::vector<int> order(size_t(ev.size()));
  for (int i = 0; i < int(order.size()); i++) order[size_t(i)] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return ev[x] < ev[y]; });
  for (int i = 0; i < int(order.size()) && i < nev; i++) {
    evals[i] = ev[order[size_t(i)]];
    for (int p = 0; p < n; p++) evecs[size_t(i) * n + p] = vec(p, order[size_t(i)]);
  }
  return nconv;
}

}  // extern "C"

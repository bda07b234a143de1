// Batched Lanczos over many patches (k_batch.hip) and the patch pipeline.
#pragma once

#include <vector>

#include "gpb_pipeline.hpp"

namespace gpb {

constexpr int kMaxPatches = 64;

// Per-patch view the batched kernels read (device memory).
struct BPatchDev {
  const float* wup;  // [40][Np] upper stencil
  const float* s;    // [Np] D^-1/2
  const float* D;    // [Np] degree
  int W, H, Wp, Hp;
  size_t Np;    // padded frame
  size_t voff;  // offset of the frame in a batched vector
  int tiles_x, tile0, ntiles, pad_;
};

struct BatchOptions {
  bool use_graph = true;  // replay restart cycles from a captured hipGraph
  int spmv_rows = 0;      // SpMV tile rows per thread (0: tuned per batch shape)
  int workgroups = 0;     // workgroups of the persistent vector kernels (0: 1024)
};

// The K smallest generalised eigenpairs of every patch's (D - W, D):
// evals[p*K + i], d_evecs[p] = [K][Np_p] (v' D v = 1, trivial pair first).
void lanczos_solve_batched(Context& c, const std::vector<Stencil>& sts, const EigOptions& opt, const BatchOptions& bo,
                           std::vector<double>& evals, const std::vector<float*>& d_evecs,
                           std::vector<LanczosStats>& stats);

}  // namespace gpb

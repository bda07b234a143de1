// Device-side plumbing shared by the gPb HIP sources: error propagation,
// owning device buffers, and the context object behind the C-ABI handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpb_common.hpp"

namespace gpb {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define GPB_HIP(call)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      throw ::gpb::HipError(std::string(#call) + " failed: " + hipGetErrorString(e_) + " (" __FILE__ ":" + \
                            std::to_string(__LINE__) + ")");                                            \
  } while (0)

#define GPB_LAUNCH_CHECK() GPB_HIP(hipGetLastError())

// Owning device allocation.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count) GPB_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  // Grow-only reservation: keeps the allocation when it is already large enough.
  void reserve(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Per-stage timing collected with HIP events on the context stream.
struct StageTimer {
  hipEvent_t ev[16][2];
  int count = 0;
  bool enabled = false;
};

struct LanczosStats {
  int iterations = 0;       // operator applications (columns)
  int restarts = 0;
  int basis = 0;            // maximum basis size m
  int block = 1;            // block width b
  int converged = 0;        // number of converged pairs
  double max_residual = 0;  // max relative generalized residual of returned pairs
  int fallback = 0;         // warm start rejected, plain solve used
  int second_passes = 0;    // Gram-Schmidt steps that needed a second pass
  int redone_steps = 0;     // optimistic-cycle steps repeated with the gated pass
  double pass2_bytes[2] = {0, 0};
};

// Per-kernel-class timing with HIP events on the context stream (enabled by
// gpb_set_profiling); durations are folded in when the stream is synchronised.
enum KernelClass {
  kKTexton = 0, kKCues, kKNonmax, kKAffinity, kKSpmv, kKDots, kKUpdate, kKRitz, kKSpb, kKGpb, kKWarm, kKUcm, kKOther,
  kNumKClasses
};
// Pinned host buffer (hipHostMalloc): copies to and from it are DMA
// transfers that stay asynchronous; a pageable buffer makes every small
// copy a staged, blocking one.
template <typename T>
struct HostBuf {
  T* p = nullptr;
  size_t n = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  void reserve(size_t count) {
    if (count <= n) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    GPB_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
};

struct KernelTimes {
  bool on = false;
  unsigned mask = ~0u;  // kernel classes that get events while on
  bool timing(int k) const { return on && ((mask >> k) & 1u); }
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> pending;  // (class, index of the start event)
  double total_ms[kNumKClasses] = {0};
  long count[kNumKClasses] = {0};
  double bytes[kNumKClasses] = {0};  // algorithmic HBM bytes of the timed launches
  hipEvent_t next() {
    if (used == pool.size()) {
      hipEvent_t e;
      GPB_HIP(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  void fold() {
    for (auto& pr : pending) {
      float ms = 0;
      GPB_HIP(hipEventElapsedTime(&ms, pool[pr.second], pool[pr.second + 1]));
      total_ms[pr.first] += ms;
      count[pr.first]++;
    }
    pending.clear();
    used = 0;
  }
  void reset() {
    for (int i = 0; i < kNumKClasses; i++) {
      total_ms[i] = 0;
      count[i] = 0;
      bytes[i] = 0;
    }
  }
};

struct Context;
void upload_tables(Context& c);
struct WarmState;  // learned warm start (k_warm.hip)

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool tables_ready = false;
  // constant tables
  DevBuf<int> schmid_off;          // member LDS offsets (texton tile)
  DevBuf<int> schmid_start;        // class starts
  DevBuf<float> schmid_w;          // [f][class]
  int schmid_classes = 0;
  // workspace (grow-only)
  std::map<std::string, std::shared_ptr<void>> scratch;
  LanczosStats last_lanczos;
  KernelTimes kt;
  double last_evals[kMaxNvec] = {0};
  // dimensions of the last whole-path run whose eigenvectors ("run.evecs")
  // and degree ("st.D") are still in the workspace; 0 once replaced
  int last_run_W = 0, last_run_H = 0, last_run_K = 0;
  double stage_ms[8] = {0};
  // eigensolver options (gpb_set_eig_options)
  int eig_K = 9;
  int eig_m = 48;
  int eig_keep = 16;
  double eig_tol = 1e-3;
  bool eig_rel = false;  // tol relative to the Ritz value (gpb_set_eig_convergence)
  int eig_max_matvecs = 20000;
  int eig_block = 1;  // 1: single-vector Lanczos; 4, 8, 12, 16: block Lanczos
  // single-vector Lanczos steps: 1 (default) partial reorthogonalisation for
  // residual targets >= 1e-4, 0 whole-basis Gram-Schmidt every step
  int reorth_mode = 2;         // 0 every step, 1 partial (gated launches), 2 partial (optimistic cycles)
  double pro_threshold = 0;    // omega trigger level, 0 = sqrt(eps_32)
  // single-vector Lanczos: replay restart cycles from captured hipGraphs
  bool lanczos_graph = true;
  // single-vector Lanczos with partial reorthogonalisation: fused steps (one
  // pass and one small kernel per step) or the separate kernels
  bool lanczos_fused = true;
  int block_fusion = 2;  // block Lanczos: 0 separate passes, 1 fused, 2 fused + DGKS-gated second local pass
  // SpMV tile rows per thread: tuned per image shape (W << 32 | H), or forced
  std::map<uint64_t, int> spmv_rows;
  std::map<uint64_t, int> spmv_band;  // tile-band width tuned with the rows (tiles)
  int spmv_rows_forced = 0;
  bool spmv_fp64 = true;   // SpMV arithmetic in FP64 (operands and results FP32 in HBM)
  bool capturing = false;
  int verbose = 0;         // > 0: solvers report each restart on stderr
  // per-patch solver statistics and eigenvalues of the last patch call
  std::vector<LanczosStats> patch_stats;
  std::vector<double> patch_evals;
  std::vector<int> patch_dims;  // w, h of each patch of the last patch call (its "bp<p>." buffers)
  int patch_K = 0;
  // learned warm start: network, optimiser and training set; the mode used
  // by the eigensolve (0 plain, 1 learned block, 2 block from a random start)
  std::shared_ptr<WarmState> warm;
  int warm_mode = 0;
  int warm_basis = 80, warm_keep = 24;  // block-solver basis and restart size
  double warm_ms = 0;  // network inference time of the last warm-started solve

  // pinned host workspace (grow-only), keyed like the device workspace
  std::map<std::string, std::shared_ptr<void>> host_scratch;
  template <typename T>
  T* host_ws(const std::string& name, size_t count) {
    auto it = host_scratch.find(name);
    if (it == host_scratch.end())
      it = host_scratch.emplace(name, std::static_pointer_cast<void>(std::make_shared<HostBuf<T>>())).first;
    auto* b = static_cast<HostBuf<T>*>(it->second.get());
    b->reserve(count);
    return b->p;
  }

  template <typename T>
  T* ws(const std::string& name, size_t count) {
    auto it = scratch.find(name);
    if (it != scratch.end()) {
      auto* b = static_cast<DevBuf<T>*>(it->second.get());
      b->reserve(count);
      return b->p;
    }
    auto b = std::make_shared<DevBuf<T>>(count);
    T* p = b->p;
    scratch[name] = std::static_pointer_cast<void>(b);
    return p;
  }
};

inline unsigned div_up(size_t a, size_t b) { return unsigned((a + b - 1) / b); }

// Fused multiply-add in the operands' own precision (kernels templated on
// their arithmetic type).
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Records a start/stop event pair around the launches in its scope.
struct KScope {
  Context& c;
  int cls;
  size_t idx = 0;
  bool active;
  KScope(Context& cc, int k, double algo_bytes) : c(cc), cls(k), active(cc.kt.timing(k)) {
    if (active) {
      c.kt.bytes[k] += algo_bytes;
      idx = c.kt.used;
      hipEvent_t a = c.kt.next();
      (void)c.kt.next();
      GPB_HIP(hipEventRecord(a, c.stream));
    }
  }
  ~KScope() {
    if (active) {
      (void)hipEventRecord(c.kt.pool[idx + 1], c.stream);
      c.kt.pending.push_back({cls, idx});
    }
  }
};

}  // namespace gpb

struct gpb_ctx {
  gpb::Context c;
};

// Whole-path orchestration on one stream (texton -> cues -> nonmax -> W ->
// eigensolve -> sPb -> gPb) and the C-ABI entry points of stages (c)-(e).
//
// Reference call chain: computeGPb (damascene/src/damascene.cu:249-661) on the
// GPU side; GlobalPb -> multiscalePb -> spectralPb (src/globalPb.cpp:15-126,
// src/multiscalePb.cpp:16-283, src/spectralPb.cpp:33-343) on the CPU side.
#include <cstring>

#include "gpb_capi.h"
#include "gpb_pipeline.hpp"

namespace gpb {

enum Stage { kStTexton = 0, kStCues, kStNonmax, kStAffinity, kStEigs, kStSpbGpb, kNumStages };

struct StageEvents {
  hipEvent_t ev[kNumStages + 1] = {};
  explicit StageEvents() {
    for (auto& e : ev) GPB_HIP(hipEventCreate(&e));
  }
  ~StageEvents() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

struct DeviceImage {
  int W, H;
  const uint32_t* rgb;
  const float* codebook;
};

struct PipelineBuffers {
  uint8_t* text;
  float* mpb_all;
  float* gcue;
  float* cues;  // optional
  float* nmax;
  float* evecs;
  float* spb;
};

static PipelineBuffers pipeline_buffers(Context& c, int W, int H, int K, bool want_cues, size_t Np) {
  const size_t N = size_t(W) * H;
  PipelineBuffers b{};
  b.text = c.ws<uint8_t>("run.text", N);
  b.mpb_all = c.ws<float>("run.mpb_all", 8 * N);
  b.gcue = c.ws<float>("run.gcue", 8 * N);
  b.cues = want_cues ? c.ws<float>("run.cues", 96 * N) : nullptr;
  b.nmax = c.ws<float>("run.nmax", N);
  b.evecs = c.ws<float>("run.evecs", size_t(K) * Np);
  b.spb = c.ws<float>("run.spb", 8 * N);
  return b;
}

void solve_eigs(Context& c, const Stencil& st, const EigOptions& eo, double* evals, float* d_evecs);

// Runs the whole path; outputs are device pointers (any may be null).
static void run_pipeline(Context& c, const DeviceImage& im, const EigOptions& eo, bool want_cues, float* d_gmax,
                         float* d_gall, float* d_orient, double* evals, PipelineBuffers* keep_buffers) {
  const int W = im.W, H = im.H;
  const size_t Np = size_t(W + 2 * kICRadius) * (H + 2 * kICRadius);
  PipelineBuffers b = pipeline_buffers(c, W, H, eo.K, want_cues, Np);
  StageEvents se;
  GPB_HIP(hipEventRecord(se.ev[0], c.stream));
  launch_texton(c, im.rgb, W, H, im.codebook, b.text);
  GPB_HIP(hipEventRecord(se.ev[1], c.stream));
  launch_local_cues(c, im.rgb, b.text, W, H, b.mpb_all, b.gcue, b.cues);
  GPB_HIP(hipEventRecord(se.ev[2], c.stream));
  launch_nonmax(c, b.mpb_all, W, H, b.nmax);
  GPB_HIP(hipEventRecord(se.ev[3], c.stream));
  Stencil st;
  build_stencil(c, b.nmax, W, H, st);
  GPB_HIP(hipEventRecord(se.ev[4], c.stream));
  solve_eigs(c, st, eo, evals, b.evecs);
  GPB_HIP(hipEventRecord(se.ev[5], c.stream));
  launch_spb(c, st, b.evecs, evals, eo.K, b.spb);
  launch_gpb(c, W, H, b.gcue, b.spb, b.nmax, d_orient, d_gmax, d_gall);
  GPB_HIP(hipEventRecord(se.ev[6], c.stream));
  GPB_HIP(hipStreamSynchronize(c.stream));
  c.kt.fold();
  for (int s = 0; s < kNumStages; s++) {
    float ms = 0;
    GPB_HIP(hipEventElapsedTime(&ms, se.ev[s], se.ev[s + 1]));
    c.stage_ms[s] = ms;
  }
  if (keep_buffers) *keep_buffers = b;
}

// Dispatch to the single-vector or the block solver (gpb_set_block_size).
// Images with fewer than 64 pixels per basis vector use the single-vector
// solver, and so does any solve whose block panel turns rank-deficient
// (stats.fallback = 2).
void solve_eigs(Context& c, const Stencil& st, const EigOptions& eo, double* evals, float* d_evecs) {
  const size_t N = size_t(st.W) * st.H;
  if (c.eig_block <= 1 || N < size_t(64) * eo.m) {
    lanczos_solve(c, st, eo, evals, d_evecs, c.last_lanczos);
    return;
  }
  BlockOptions bo;
  bo.block = c.eig_block;
  try {
    block_lanczos_solve(c, st, eo, bo, evals, d_evecs, c.last_lanczos);
  } catch (const BlockBreakdown&) {
    const int spent = c.last_lanczos.iterations;
    lanczos_solve(c, st, eo, evals, d_evecs, c.last_lanczos);
    c.last_lanczos.fallback = 2;
    c.last_lanczos.iterations += spent;
  }
}

static EigOptions ctx_eig_options(const Context& c) {
  EigOptions e;
  e.K = c.eig_K;
  e.m = c.eig_m;
  e.keep = c.eig_keep;
  e.tol = c.eig_tol;
  e.max_matvecs = c.eig_max_matvecs;
  return e;
}

}  // namespace gpb

using namespace gpb;

namespace {

template <typename F>
int guarded2(gpb_ctx* ctx, F&& fn) {
  if (!ctx) return GPB_ERR_ARG;
  try {
    GPB_HIP(hipSetDevice(ctx->c.device));
    fn();
    return GPB_OK;
  } catch (const ArgError& e) {
    ctx->c.err = e.what();
    return GPB_ERR_ARG;
  } catch (const HipError& e) {
    ctx->c.err = e.what();
    return GPB_ERR_HIP;
  } catch (const NoConvergence& e) {
    ctx->c.err = e.what();
    return GPB_ERR_NOCONV;
  } catch (const std::exception& e) {
    ctx->c.err = e.what();
    return GPB_ERR_INTERNAL;
  }
}

void check_dims(int width, int height) {
  if (width <= 0 || height <= 0) throw ArgError("image dimensions must be positive");
  if (size_t(width) * size_t(height) > (size_t(1) << 31)) throw ArgError("image too large");
}

// Copy a padded [K][Np] device block into an unpadded [K][H][W] host array.
void unpad_to_host(Context& c, const float* d, int K, int W, int H, float* host) {
  const int Wp = W + 2 * kICRadius;
  const size_t Np = size_t(Wp) * (H + 2 * kICRadius);
  for (int k = 0; k < K; k++)
    GPB_HIP(hipMemcpy2DAsync(host + size_t(k) * W * H, size_t(W) * 4, d + size_t(k) * Np + size_t(kICRadius) * Wp + kICRadius,
                             size_t(Wp) * 4, size_t(W) * 4, H, hipMemcpyDeviceToHost, c.stream));
}

}  // namespace

extern "C" {

int gpb_set_eig_options(gpb_ctx* ctx, int nvec, int basis, int keep, double tol, int max_matvecs) {
  return guarded2(ctx, [&] {
    if (nvec < 2 || nvec > kMaxNvec || basis < nvec + 2 || basis > 320 || keep < nvec || keep >= basis || !(tol > 0) ||
        max_matvecs < basis)
      throw ArgError("invalid eigensolver options");
    ctx->c.eig_K = nvec;
    ctx->c.eig_m = basis;
    ctx->c.eig_keep = keep;
    ctx->c.eig_tol = tol;
    ctx->c.eig_max_matvecs = max_matvecs;
  });
}

int gpb_set_spmv_tile(gpb_ctx* ctx, int rows) {
  return guarded2(ctx, [&] {
    if (!(rows == 0 || rows == 2 || rows == 4 || rows == 8)) throw ArgError("SpMV tile rows must be 0, 2, 4 or 8");
    ctx->c.spmv_rows_forced = rows;
  });
}

int gpb_get_spmv_tile(gpb_ctx* ctx, int width, int height, int* rows) {
  return guarded2(ctx, [&] {
    if (!rows) throw ArgError("null pointer");
    if (ctx->c.spmv_rows_forced) {
      *rows = ctx->c.spmv_rows_forced;
      return;
    }
    auto it = ctx->c.spmv_rows.find((uint64_t(width) << 32) | uint32_t(height));
    *rows = it == ctx->c.spmv_rows.end() ? 0 : it->second;
  });
}

int gpb_set_block_size(gpb_ctx* ctx, int block) {
  return guarded2(ctx, [&] {
    if (!(block == 1 || block == 4 || block == 8 || block == 12 || block == 16))
      throw ArgError("block size must be 1, 4, 8, 12 or 16");
    ctx->c.eig_block = block;
  });
}

int gpb_set_profiling(gpb_ctx* ctx, int on) {
  return guarded2(ctx, [&] {
    GPB_HIP(hipStreamSynchronize(ctx->c.stream));
    ctx->c.kt.fold();
    ctx->c.kt.on = on != 0;
    ctx->c.kt.reset();
  });
}

int gpb_kernel_time(gpb_ctx* ctx, int kernel_class, double* total_ms, long* launches, double* algo_bytes) {
  return guarded2(ctx, [&] {
    if (kernel_class < 0 || kernel_class >= kNumKClasses || !total_ms || !launches) throw ArgError("bad kernel class");
    GPB_HIP(hipStreamSynchronize(ctx->c.stream));
    ctx->c.kt.fold();
    *total_ms = ctx->c.kt.total_ms[kernel_class];
    *launches = ctx->c.kt.count[kernel_class];
    if (algo_bytes) *algo_bytes = ctx->c.kt.bytes[kernel_class];
  });
}

int gpb_get_stats(gpb_ctx* ctx, gpb_stats* out) {
  return guarded2(ctx, [&] {
    if (!out) throw ArgError("null pointer");
    std::memset(out, 0, sizeof(*out));
    for (int s = 0; s < 6; s++) out->stage_ms[s] = ctx->c.stage_ms[s];
    out->lanczos_matvecs = ctx->c.last_lanczos.iterations;
    out->lanczos_restarts = ctx->c.last_lanczos.restarts;
    out->lanczos_basis = ctx->c.last_lanczos.basis;
    out->lanczos_block = ctx->c.last_lanczos.block;
    out->converged = ctx->c.last_lanczos.converged;
    out->max_residual = ctx->c.last_lanczos.max_residual;
    out->fallback = ctx->c.last_lanczos.fallback;
    out->reorth_second_passes = ctx->c.last_lanczos.second_passes;
    for (int i = 0; i < kMaxNvec; i++) out->evals[i] = ctx->c.last_evals[i];
  });
}

int gpb_build_w(gpb_ctx* ctx, int width, int height, const float* mpb_nmax, float* w_upper, float* degree) {
  return guarded2(ctx, [&] {
    check_dims(width, height);
    if (!mpb_nmax || !w_upper || !degree) throw ArgError("null pointer");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    float* d_m = c.ws<float>("io.nmax", N);
    GPB_HIP(hipMemcpyAsync(d_m, mpb_nmax, N * 4, hipMemcpyHostToDevice, c.stream));
    Stencil st;
    build_stencil(c, d_m, width, height, st);
    unpad_to_host(c, st.wup, kStencilHalf, width, height, w_upper);
    unpad_to_host(c, st.D, 1, width, height, degree);
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

int gpb_eigs(gpb_ctx* ctx, int width, int height, const float* mpb_nmax, double* evals, float* evecs) {
  return guarded2(ctx, [&] {
    check_dims(width, height);
    if (!mpb_nmax || !evals) throw ArgError("null pointer");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    const EigOptions eo = ctx_eig_options(c);
    float* d_m = c.ws<float>("io.nmax", N);
    GPB_HIP(hipMemcpyAsync(d_m, mpb_nmax, N * 4, hipMemcpyHostToDevice, c.stream));
    Stencil st;
    build_stencil(c, d_m, width, height, st);
    float* d_ev = c.ws<float>("io.evecs", size_t(eo.K) * st.Np);
    solve_eigs(c, st, eo, evals, d_ev);
    for (int i = 0; i < eo.K; i++) c.last_evals[i] = evals[i];
    if (evecs) unpad_to_host(c, d_ev, eo.K, width, height, evecs);
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

int gpb_spectral_pb(gpb_ctx* ctx, int width, int height, const float* mpb_nmax, float* spb, double* evals) {
  return guarded2(ctx, [&] {
    check_dims(width, height);
    if (!mpb_nmax || !spb) throw ArgError("null pointer");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    const EigOptions eo = ctx_eig_options(c);
    float* d_m = c.ws<float>("io.nmax", N);
    GPB_HIP(hipMemcpyAsync(d_m, mpb_nmax, N * 4, hipMemcpyHostToDevice, c.stream));
    Stencil st;
    build_stencil(c, d_m, width, height, st);
    float* d_ev = c.ws<float>("io.evecs", size_t(eo.K) * st.Np);
    double ev[kMaxNvec];
    solve_eigs(c, st, eo, ev, d_ev);
    for (int i = 0; i < eo.K; i++) c.last_evals[i] = ev[i];
    float* d_spb = c.ws<float>("io.spb", 8 * N);
    launch_spb(c, st, d_ev, ev, eo.K, d_spb);
    GPB_HIP(hipMemcpyAsync(spb, d_spb, 8 * N * 4, hipMemcpyDeviceToHost, c.stream));
    if (evals) std::memcpy(evals, ev, sizeof(double) * eo.K);
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

int gpb_spb_from_evecs(gpb_ctx* ctx, int width, int height, int nvec, const float* evecs, const double* evals,
                       float* spb) {
  return guarded2(ctx, [&] {
    check_dims(width, height);
    if (!evecs || !evals || !spb || nvec < 2 || nvec > kMaxNvec) throw ArgError("bad arguments");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    Stencil st;
    st.W = width;
    st.H = height;
    st.Wp = width + 2 * kICRadius;
    st.Hp = height + 2 * kICRadius;
    st.Np = size_t(st.Wp) * st.Hp;
    float* d_ev = c.ws<float>("io.evecs", size_t(nvec) * st.Np);
    for (int k = 0; k < nvec; k++)
      GPB_HIP(hipMemcpy2DAsync(d_ev + size_t(k) * st.Np + st.pad_index(0, 0), size_t(st.Wp) * 4, evecs + size_t(k) * N,
                               size_t(width) * 4, size_t(width) * 4, height, hipMemcpyHostToDevice, c.stream));
    float* d_spb = c.ws<float>("io.spb", 8 * N);
    launch_spb(c, st, d_ev, evals, nvec, d_spb);
    GPB_HIP(hipMemcpyAsync(spb, d_spb, 8 * N * 4, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

int gpb_compute_device(gpb_ctx* ctx, int width, int height, const uint32_t* d_packed, const float* d_codebook,
                       float* d_gpb_max, float* d_gpb_all, float* d_gpb_orient) {
  return guarded2(ctx, [&] {
    check_dims(width, height);
    if (!d_packed || !d_codebook) throw ArgError("null pointer");
    Context& c = ctx->c;
    const EigOptions eo = ctx_eig_options(c);
    double ev[kMaxNvec];
    run_pipeline(c, DeviceImage{width, height, d_packed, d_codebook}, eo, false, d_gpb_max, d_gpb_all, d_gpb_orient, ev,
                 nullptr);
    for (int i = 0; i < eo.K; i++) c.last_evals[i] = ev[i];
  });
}

int gpb_compute(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook, float* gpb_max,
                float* gpb_all, float* gpb_orient, float* mpb_nmax) {
  return guarded2(ctx, [&] {
    check_dims(width, height);
    if (!packed || !codebook) throw ArgError("null pointer");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    const EigOptions eo = ctx_eig_options(c);
    uint32_t* d_rgb = c.ws<uint32_t>("io.rgb", N);
    float* d_cb = c.ws<float>("io.codebook", kSchmidNF * kTextonK);
    float* d_gmax = gpb_max ? c.ws<float>("io.gmax", N) : nullptr;
    float* d_gall = gpb_all ? c.ws<float>("io.gall", 8 * N) : nullptr;
    float* d_or = gpb_orient ? c.ws<float>("io.gorient", 8 * N) : nullptr;
    GPB_HIP(hipMemcpyAsync(d_rgb, packed, N * 4, hipMemcpyHostToDevice, c.stream));
    GPB_HIP(hipMemcpyAsync(d_cb, codebook, kSchmidNF * kTextonK * 4, hipMemcpyHostToDevice, c.stream));
    double ev[kMaxNvec];
    PipelineBuffers b;
    run_pipeline(c, DeviceImage{width, height, d_rgb, d_cb}, eo, false, d_gmax, d_gall, d_or, ev, &b);
    for (int i = 0; i < eo.K; i++) c.last_evals[i] = ev[i];
    if (gpb_max) GPB_HIP(hipMemcpyAsync(gpb_max, d_gmax, N * 4, hipMemcpyDeviceToHost, c.stream));
    if (gpb_all) GPB_HIP(hipMemcpyAsync(gpb_all, d_gall, 8 * N * 4, hipMemcpyDeviceToHost, c.stream));
    if (gpb_orient) GPB_HIP(hipMemcpyAsync(gpb_orient, d_or, 8 * N * 4, hipMemcpyDeviceToHost, c.stream));
    if (mpb_nmax) GPB_HIP(hipMemcpyAsync(mpb_nmax, b.nmax, N * 4, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

}  // extern "C"

extern "C" int gpb_apply_laplacian(gpb_ctx* ctx, int width, int height, const float* mpb_nmax, int ncols, int block,
                                   const float* x, float* y) {
  return guarded2(ctx, [&] {
    check_dims(width, height);
    if (!mpb_nmax || !x || !y || ncols < 1 || ncols > 16) throw ArgError("bad arguments");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    float* d_m = c.ws<float>("io.nmax", N);
    GPB_HIP(hipMemcpyAsync(d_m, mpb_nmax, N * 4, hipMemcpyHostToDevice, c.stream));
    Stencil st;
    build_stencil(c, d_m, width, height, st);
    float* dx = c.ws<float>("io.lx", size_t(ncols) * st.Np);
    float* dy = c.ws<float>("io.ly", size_t(ncols) * st.Np);
    launch_fill(c, dx, size_t(ncols) * st.Np, 0.f);
    launch_fill(c, dy, size_t(ncols) * st.Np, 0.f);
    for (int k = 0; k < ncols; k++)
      GPB_HIP(hipMemcpy2DAsync(dx + size_t(k) * st.Np + st.pad_index(0, 0), size_t(st.Wp) * 4, x + size_t(k) * N,
                               size_t(width) * 4, size_t(width) * 4, height, hipMemcpyHostToDevice, c.stream));
    if (block == 1) {
      for (int k = 0; k < ncols; k++) debug_block_spmv(c, st, 1, dx + size_t(k) * st.Np, dy + size_t(k) * st.Np);
    } else {
      if (block != ncols) throw ArgError("block must equal ncols");
      debug_block_spmv(c, st, block, dx, dy);
    }
    unpad_to_host(c, dy, ncols, width, height, y);
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

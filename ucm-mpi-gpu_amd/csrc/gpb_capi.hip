// C-ABI entry points of libgpb.so.  Each wraps a C++ implementation in an
// exception guard that turns failures into a status code and a message on the
// context (the reference throws int64_t(1) after cudaDeviceReset instead,
// damascene/include/exception.h:16-28).
#include <cstring>

#include "gpb_capi.h"
#include "gpb_device.hpp"
#include "gpb_pipeline.hpp"

using namespace gpb;

namespace {

template <typename F>
int guarded(gpb_ctx* ctx, F&& fn) {
  try {
    if (ctx) GPB_HIP(hipSetDevice(ctx->c.device));
    fn();
    return GPB_OK;
  } catch (const ArgError& e) {
    if (ctx) ctx->c.err = e.what();
    return GPB_ERR_ARG;
  } catch (const HipError& e) {
    if (ctx) ctx->c.err = e.what();
    return GPB_ERR_HIP;
  } catch (const NoConvergence& e) {
    if (ctx) ctx->c.err = e.what();
    return GPB_ERR_NOCONV;
  } catch (const std::exception& e) {
    if (ctx) ctx->c.err = e.what();
    return GPB_ERR_INTERNAL;
  }
}

void check_image(int width, int height) {
  if (width <= 0 || height <= 0) throw ArgError("image dimensions must be positive");
  if (size_t(width) * size_t(height) > (size_t(1) << 31)) throw ArgError("image too large (more than 2^31 pixels)");
}

}  // namespace

extern "C" {

int gpb_ctx_create(int device, gpb_ctx** out) {
  if (!out) return GPB_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GPB_ERR_HIP;
  auto* ctx = new gpb_ctx();
  ctx->c.device = device;
  const int rc = guarded(ctx, [&] {
    GPB_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    upload_tables(ctx->c);
    // this device's address of the no-allocation entry point's scratch array
    if (noalloc_scratch_setup(device) != GPB_OK) throw HipError("hipGetSymbolAddress(scratch) failed");
  });
  if (rc != GPB_OK) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return GPB_OK;
}

void gpb_ctx_destroy(gpb_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  ctx->c.scratch.clear();
  if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

const char* gpb_ctx_last_error(const gpb_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int gpb_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void gpb_pack_rgb(const uint8_t* rgb, size_t npixels, uint32_t* packed) {
  for (size_t i = 0; i < npixels; i++) packed[i] = pack_rgb(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
}

int gpb_synth_voronoi(int width, int height, int regions, double noise_sigma, uint64_t seed, int classes,
                      uint8_t* rgb, int32_t* region) {
  return guarded(nullptr, [&] {
    check_image(width, height);
    if (!rgb || regions <= 0 || noise_sigma < 0 || (classes != 1 && classes != 3)) throw ArgError("bad synth arguments");
    SynthParams p;
    p.width = width;
    p.height = height;
    p.regions = regions;
    p.noise_sigma = noise_sigma;
    p.seed = seed;
    p.classes = classes;
    synth_voronoi(p, rgb, region);
  });
}

int gpb_make_codebook(const uint32_t* packed, int width, int height, uint64_t seed, int iters, int max_samples,
                      float* codebook) {
  return guarded(nullptr, [&] {
    check_image(width, height);
    if (!packed || !codebook || iters < 0 || max_samples < kTextonK) throw ArgError("bad codebook arguments");
    if (size_t(width) * height < size_t(kTextonK)) throw ArgError("image smaller than the codebook");
    make_codebook(packed, width, height, seed, iters, max_samples, codebook);
  });
}

int gpb_texton(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook, uint8_t* labels) {
  if (!ctx) return GPB_ERR_ARG;
  return guarded(ctx, [&] {
    check_image(width, height);
    if (!packed || !codebook || !labels) throw ArgError("null pointer");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    uint32_t* d_rgb = c.ws<uint32_t>("io.rgb", N);
    float* d_cb = c.ws<float>("io.codebook", kSchmidNF * kTextonK);
    uint8_t* d_lab = c.ws<uint8_t>("io.labels", N);
    GPB_HIP(hipMemcpyAsync(d_rgb, packed, N * 4, hipMemcpyHostToDevice, c.stream));
    GPB_HIP(hipMemcpyAsync(d_cb, codebook, kSchmidNF * kTextonK * 4, hipMemcpyHostToDevice, c.stream));
    launch_texton(c, d_rgb, width, height, d_cb, d_lab);
    GPB_HIP(hipMemcpyAsync(labels, d_lab, N, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

int gpb_texton_features(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook,
                        const int32_t* superpixels, int nlabels, double* descr) {
  if (!ctx) return GPB_ERR_ARG;
  return guarded(ctx, [&] {
    check_image(width, height);
    if (!packed || !codebook || !superpixels || !descr || nlabels < 1) throw ArgError("null pointer or no labels");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    uint32_t* d_rgb = c.ws<uint32_t>("io.rgb", N);
    float* d_cb = c.ws<float>("io.codebook", kSchmidNF * kTextonK);
    uint8_t* d_lab = c.ws<uint8_t>("io.labels", N);
    int32_t* d_sp = c.ws<int32_t>("io.superpixels", N);
    double* d_descr = c.ws<double>("io.descr", size_t(nlabels) * kTextonK);
    GPB_HIP(hipMemcpyAsync(d_rgb, packed, N * 4, hipMemcpyHostToDevice, c.stream));
    GPB_HIP(hipMemcpyAsync(d_cb, codebook, kSchmidNF * kTextonK * 4, hipMemcpyHostToDevice, c.stream));
    GPB_HIP(hipMemcpyAsync(d_sp, superpixels, N * 4, hipMemcpyHostToDevice, c.stream));
    launch_texton(c, d_rgb, width, height, d_cb, d_lab);
    launch_texton_features(c, d_lab, width, height, d_sp, nlabels, d_descr);
    GPB_HIP(hipMemcpyAsync(descr, d_descr, sizeof(double) * size_t(nlabels) * kTextonK, hipMemcpyDeviceToHost,
                           c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

int gpb_texton_patch_features(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook,
                              int ncentres, const int32_t* centres, double* descr) {
  if (!ctx) return GPB_ERR_ARG;
  return guarded(ctx, [&] {
    check_image(width, height);
    if (!packed || !codebook || ncentres < 0 || (ncentres > 0 && (!centres || !descr))) throw ArgError("null pointer");
    for (int k = 0; k < ncentres; k++)
      if (centres[2 * k] < 0 || centres[2 * k] >= height || centres[2 * k + 1] < 0 || centres[2 * k + 1] >= width)
        throw ArgError("centre outside the image");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    uint32_t* d_rgb = c.ws<uint32_t>("io.rgb", N);
    float* d_cb = c.ws<float>("io.codebook", kSchmidNF * kTextonK);
    uint8_t* d_lab = c.ws<uint8_t>("io.labels", N);
    GPB_HIP(hipMemcpyAsync(d_rgb, packed, N * 4, hipMemcpyHostToDevice, c.stream));
    GPB_HIP(hipMemcpyAsync(d_cb, codebook, kSchmidNF * kTextonK * 4, hipMemcpyHostToDevice, c.stream));
    launch_texton(c, d_rgb, width, height, d_cb, d_lab);
    if (ncentres > 0) {
      int32_t* d_ctr = c.ws<int32_t>("io.centres", size_t(2) * ncentres);
      double* d_descr = c.ws<double>("io.descr", size_t(ncentres) * kTextonK);
      GPB_HIP(hipMemcpyAsync(d_ctr, centres, sizeof(int32_t) * 2 * size_t(ncentres), hipMemcpyHostToDevice, c.stream));
      launch_texton_patch_features(c, d_lab, width, height, d_ctr, ncentres, d_descr);
      GPB_HIP(hipMemcpyAsync(descr, d_descr, sizeof(double) * size_t(ncentres) * kTextonK, hipMemcpyDeviceToHost,
                             c.stream));
    }
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

}  // extern "C"

// ---- stage (b) ------------------------------------------------------------------
extern "C" {

int gpb_lab_quantize(gpb_ctx* ctx, int width, int height, const uint32_t* packed, uint8_t* bins, int* nbins) {
  if (!ctx) return GPB_ERR_ARG;
  return guarded(ctx, [&] {
    check_image(width, height);
    if (!packed || !bins) throw ArgError("null pointer");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    uint32_t* d_rgb = c.ws<uint32_t>("io.rgb", N);
    uint8_t* d_q = c.ws<uint8_t>("io.q", 3 * N);
    int* d_nb = c.ws<int>("io.nb", 3);
    GPB_HIP(hipMemcpyAsync(d_rgb, packed, N * 4, hipMemcpyHostToDevice, c.stream));
    launch_lab_quantize(c, d_rgb, N, d_q, d_nb);
    GPB_HIP(hipMemcpyAsync(bins, d_q, 3 * N, hipMemcpyDeviceToHost, c.stream));
    if (nbins) GPB_HIP(hipMemcpyAsync(nbins, d_nb, 3 * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

int gpb_local_cues(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const uint8_t* textons, float* cues,
                   float* mpb_all) {
  if (!ctx) return GPB_ERR_ARG;
  return guarded(ctx, [&] {
    check_image(width, height);
    if (!packed || !textons || !mpb_all) throw ArgError("null pointer");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    uint32_t* d_rgb = c.ws<uint32_t>("io.rgb", N);
    uint8_t* d_t = c.ws<uint8_t>("io.labels", N);
    float* d_m = c.ws<float>("io.mpb_all", 8 * N);
    float* d_g = c.ws<float>("io.gcue", 8 * N);
    float* d_c = cues ? c.ws<float>("io.cues", 96 * N) : nullptr;
    GPB_HIP(hipMemcpyAsync(d_rgb, packed, N * 4, hipMemcpyHostToDevice, c.stream));
    GPB_HIP(hipMemcpyAsync(d_t, textons, N, hipMemcpyHostToDevice, c.stream));
    launch_local_cues(c, d_rgb, d_t, width, height, d_m, d_g, d_c);
    GPB_HIP(hipMemcpyAsync(mpb_all, d_m, 8 * N * 4, hipMemcpyDeviceToHost, c.stream));
    if (cues) GPB_HIP(hipMemcpyAsync(cues, d_c, 96 * N * 4, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

int gpb_nonmax(gpb_ctx* ctx, int width, int height, const float* mpb_all, float* nmax) {
  if (!ctx) return GPB_ERR_ARG;
  return guarded(ctx, [&] {
    check_image(width, height);
    if (!mpb_all || !nmax) throw ArgError("null pointer");
    Context& c = ctx->c;
    const size_t N = size_t(width) * height;
    float* d_m = c.ws<float>("io.mpb_all", 8 * N);
    float* d_n = c.ws<float>("io.nmax", N);
    GPB_HIP(hipMemcpyAsync(d_m, mpb_all, 8 * N * 4, hipMemcpyHostToDevice, c.stream));
    launch_nonmax(c, d_m, width, height, d_n);
    GPB_HIP(hipMemcpyAsync(nmax, d_n, N * 4, hipMemcpyDeviceToHost, c.stream));
    GPB_HIP(hipStreamSynchronize(c.stream));
  });
}

}  // extern "C"

// ---- diagnostics ------------------------------------------------------------------
extern "C" int gpb_ic_program(int left, int right, int top, int bottom, int8_t* conn, int8_t* sites, int max_sites) {
  return guarded(nullptr, [&] {
    if (left < 0 || left > 5 || right < 0 || right > 5 || top < 0 || top > 5 || bottom < 0 || bottom > 5 || !conn)
      throw ArgError("bad program arguments");
    const ICProgram& p = ic_program(left, right, top, bottom);
    int ddx[kStencilFull], ddy[kStencilFull];
    disc_offsets(ddx, ddy);
    std::memset(conn, 0, kStencilFull * 5);
    for (const ICConnection& c : p.conn) {
      for (int k = 0; k < kStencilFull; k++)
        if (ddx[k] == c.dx && ddy[k] == c.dy) {
          int8_t* e = conn + 5 * k;
          e[0] = 1;  // valid
          e[1] = int8_t(c.found);
          e[2] = int8_t(c.site_count);
          e[3] = int8_t(c.site_start & 0x7f);
          e[4] = int8_t(c.site_start >> 7);
        }
    }
    if (sites) {
      if (int(p.site_dx.size()) > max_sites) throw ArgError("site buffer too small");
      for (size_t i = 0; i < p.site_dx.size(); i++) {
        sites[2 * i] = p.site_dx[i];
        sites[2 * i + 1] = p.site_dy[i];
      }
    }
  });
}

// gpb_nonmax_noalloc: the oriented nonmax of gpb_nonmax on device buffers,
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  Its two
// kernels keep their per-pixel scratch -- the largest oriented mPb (FP32) and
// its orientation (one byte) -- in the file-scope device array `scratch`
// below instead of allocated memory.  Every GPU the library runs on holds its
// own copy of the array, at its own address, until the process ends.
#include <mutex>

#include "gpb_capi.h"
#include "gpb_pipeline.hpp"

namespace {

constexpr size_t kMaxPixels = size_t(GPB_NOALLOC_MAX_PIXELS);
// a[kMaxPixels] floats, then k[kMaxPixels] bytes: 20 MiB
constexpr size_t kScratchBytes = kMaxPixels * (sizeof(float) + 1);

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
// Device address of each device's `scratch` (hipGetSymbolAddress with that
// device current).  In host code the name `scratch` is a host-side stand-in,
// not a device address: only these pointers reach kernels.
void* g_scratch_dev[kMaxDevices] = {};
std::mutex g_mu;

}  // namespace

namespace gpb {

int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return GPB_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_dev[device]) return GPB_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return GPB_ERR_HIP;
  g_scratch_dev[device] = p;
  return GPB_OK;
}

}  // namespace gpb

extern "C" int gpb_nonmax_noalloc(int width, int height, const float* d_mpb_all, float* d_nmax, void* stream) {
  if (width < 1 || height < 1 || !d_mpb_all || !d_nmax) return GPB_ERR_ARG;
  if (height > 65535) return GPB_ERR_ARG;  // one grid row per image row
  if (size_t(width) * size_t(height) > kMaxPixels) return GPB_ERR_TOO_LARGE;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return GPB_ERR_HIP;
  if (!g_scratch_dev[dev]) {
    const int rc = gpb::noalloc_scratch_setup(dev);
    if (rc != GPB_OK) return rc;
  }
  auto* base = static_cast<unsigned char*>(g_scratch_dev[dev]);
  float* a = reinterpret_cast<float*>(base);
  uint8_t* k = base + kMaxPixels * sizeof(float);
  try {
    gpb::launch_nonmax_scratch(static_cast<hipStream_t>(stream), d_mpb_all, width, height, a, k, d_nmax);
  } catch (const std::exception&) {
    return GPB_ERR_HIP;
  }
  return GPB_OK;
}

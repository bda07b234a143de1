// Byte framing of the reference's matrix messages, for moving gPb outputs
// between processes or nodes over transports other than MPI.
//
// The reference sends a matrix as two MPI messages, a header of four ints
// {rows, cols, channels, OpenCV type} and then rows * cols * channels elements
// of the type's depth (MpiSendMat / MpiRecvMat, src/utilities.cpp:382-413);
// a vector of matrices is preceded by one int, the count
// (MpiSendMatVec / MpiRecvMatVec, :415-460).  A frame here is those messages
// back to back in one buffer: [count] then, per matrix, the 16-byte header and
// the payload, all in host byte order with no padding.
#include <cstdint>
#include <cstring>

#include "gpb_capi.h"

namespace {

// Element size of an OpenCV depth (type & 7): 8U, 8S, 16U, 16S, 32S, 32F, 64F.
int depth_bytes(int type) {
  static const int sz[8] = {1, 1, 2, 2, 4, 4, 8, 0};
  return type < 0 ? 0 : sz[type & 7];
}

bool payload_bytes(const int* d, size_t* out) {
  const int rows = d[0], cols = d[1], ch = d[2], type = d[3];
  const int es = depth_bytes(type);
  // OpenCV's type code carries the channel count too: type = depth + 8 (channels - 1)
  if (rows < 0 || cols < 0 || ch < 1 || ch > 512 || es == 0 || ((type >> 3) + 1) != ch) return false;
  *out = size_t(rows) * size_t(cols) * size_t(ch) * size_t(es);
  return true;
}

}  // namespace

extern "C" {

size_t gpb_frame_size(int count, const int* dims) {
  if (count < 0 || (count > 0 && !dims)) return 0;
  size_t n = sizeof(int32_t);
  for (int i = 0; i < count; i++) {
    size_t b = 0;
    if (!payload_bytes(dims + 4 * i, &b)) return 0;
    n += 4 * sizeof(int32_t) + b;
  }
  return n;
}

int gpb_frame_mats(int count, const int* dims, const void* const* data, void* out, size_t out_size) {
  const size_t need = gpb_frame_size(count, dims);
  if (need == 0 || !out || out_size < need || (count > 0 && !data)) return GPB_ERR_ARG;
  auto* p = static_cast<unsigned char*>(out);
  const int32_t c = count;
  std::memcpy(p, &c, sizeof c);
  p += sizeof c;
  for (int i = 0; i < count; i++) {
    const int32_t h[4] = {dims[4 * i], dims[4 * i + 1], dims[4 * i + 2], dims[4 * i + 3]};
    std::memcpy(p, h, sizeof h);
    p += sizeof h;
    size_t b = 0;
    payload_bytes(dims + 4 * i, &b);
    if (b && !data[i]) return GPB_ERR_ARG;
    if (b) std::memcpy(p, data[i], b);
    p += b;
  }
  return GPB_OK;
}

int gpb_unframe_mats(const void* buf, size_t size, int max_count, int* count, int* dims, size_t* offsets) {
  if (!buf || !count || size < sizeof(int32_t)) return GPB_ERR_ARG;
  const auto* p = static_cast<const unsigned char*>(buf);
  int32_t c = 0;
  std::memcpy(&c, p, sizeof c);
  if (c < 0) return GPB_ERR_ARG;
  *count = c;
  if (c > max_count || (c > 0 && (!dims || !offsets))) return GPB_ERR_ARG;
  size_t at = sizeof c;
  for (int i = 0; i < c; i++) {
    if (size - at < 4 * sizeof(int32_t)) return GPB_ERR_ARG;
    int32_t h[4];
    std::memcpy(h, p + at, sizeof h);
    at += sizeof h;
    const int d[4] = {h[0], h[1], h[2], h[3]};
    size_t b = 0;
    if (!payload_bytes(d, &b) || size - at < b) return GPB_ERR_ARG;
    for (int k = 0; k < 4; k++) dims[4 * i + k] = d[k];
    offsets[i] = at;
    at += b;
  }
  return at == size ? GPB_OK : GPB_ERR_ARG;
}

}  // extern "C"

// Host-side copy between a caller's (pageable) buffer and a pinned staging
// chunk, spread over the host's cores: one thread moves about 10 GB/s, so a
// single-threaded copy would bound the host-buffer entry points' transfers.
#include <algorithm>
#include <cstddef>
#include <cstring>

namespace gpb {

void parallel_copy(void* dst, const void* src, size_t bytes) {
  constexpr size_t kPiece = size_t(1) << 20;
  const long pieces = long((bytes + kPiece - 1) / kPiece);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < pieces; i++) {
    const size_t off = size_t(i) * kPiece;
    std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, std::min(kPiece, bytes - off));
  }
}

}  // namespace gpb

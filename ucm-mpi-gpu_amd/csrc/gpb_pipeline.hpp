// Internal launch interface between the C-ABI layer and the stage kernels.
#pragma once

#include "gpb_device.hpp"

namespace gpb {

struct NoConvergence : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// stage (a)
void upload_schmid(Context& c);
void launch_texton(Context& c, const uint32_t* d_rgb, int W, int H, const float* d_codebook, uint8_t* d_labels);

}  // namespace gpb

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

// stage (b)
void launch_lab_quantize(Context& c, const uint32_t* d_rgb, size_t N, uint8_t* d_q /* 3N */, int* d_nb /* 3 */);
void launch_local_cues(Context& c, const uint32_t* d_rgb, const uint8_t* d_text, int W, int H, float* d_mpb_all,
                       float* d_gcue, float* d_cues /* optional [12][8][N] */);
void launch_nonmax(Context& c, const float* d_mpb_all, int W, int H, float* d_nmax);

}  // namespace gpb

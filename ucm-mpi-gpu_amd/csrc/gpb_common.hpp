// Host-side definitions shared by the gPb library sources: fixed pipeline
// constants, the packed-pixel format, border reflection and the tables that the
// HIP kernels read (Schmid bank, orientation slices, smoothing kernels, nonmax
// and sPb filters, intervening-contour walk programs).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gpb {

// ---- pipeline constants -------------------------------------------------
constexpr int kNumOri = 8;          // orientations (mex_pb_parts_final_selected.cpp:79)
constexpr int kNumSlices = 2 * kNumOri;
constexpr int kSchmidNF = 13;       // Schmid bank size (SPFeatures.cpp:188)
constexpr int kSchmidSup = 49;      // Schmid support (SPFeatures.cpp:188)
constexpr int kSchmidHalf = kSchmidSup / 2;
constexpr int kTextonK = 32;        // codebook columns (SPFeatures.cpp:44)
constexpr int kCueBorder = 30;      // mirror border for local cues (mex_pb_parts_final_selected.cpp:86)
constexpr int kLabBins = 25;        // L, a, b quantisation bins (:80-82)
constexpr int kICRadius = 5;        // intervening-contour radius (buildW.cpp:24)
constexpr int kStencilHalf = 40;    // disc offsets with (dy>0) or (dy==0, dx>0) at radius 5
constexpr int kStencilFull = 81;    // all disc offsets including the centre
constexpr int kNumCues = 12;        // bg1..3, cga1..3, cgb1..3, tg1..3
constexpr int kNumXcd = 8;         // accelerator complex dies of an MI355X (one L2 each)
constexpr int kMaxNvec = 32;        // eigenpairs per solve (spectralPb's CPU default is 17)

// Truncated pi used by the reference for filter phases and nonmax angles
// (include/UCM_CONSTANTS.h:4).
constexpr double kRefPI = 3.1415926535;

// mPb cue weights (multiscalePb.cpp:20) and gPb weights (globalPb.cpp:32),
// cue order bg1 bg2 bg3 cga1 cga2 cga3 cgb1 cgb2 cgb3 tg1 tg2 tg3 [sPb].
extern const double kMPbWeights[kNumCues];
extern const double kGPbWeights[kNumCues + 1];

// Radii per channel (mex_pb_parts_final_selected.cpp:90-92).
extern const int kRadiusBg[3];
extern const int kRadiusCg[3];

// ---- packed pixels ------------------------------------------------------
// Packed pixel: R in bits 0-7, G in 8-15, B in 16-23 (globalPb.cpp:264-276).
inline uint32_t pack_rgb(uint8_t r, uint8_t g, uint8_t b) {
  return uint32_t(r) | (uint32_t(g) << 8) | (uint32_t(b) << 16);
}

// OpenCV CV_BGR2GRAY for 8-bit input: fixed point, 14 fractional bits,
// round half up (SPFeatures.cpp:27 calls it).
inline int grey_u8(uint32_t px) {
  const int r = px & 0xff, g = (px >> 8) & 0xff, b = (px >> 16) & 0xff;
  return (r * 4899 + g * 9617 + b * 1868 + 8192) >> 14;
}

// Half-sample symmetric index (OpenCV BORDER_REFLECT, lib_image border_mirror_2D),
// extended periodically so that any integer maps into [0, n).
inline int reflect_index(int i, int n) {
  if (n == 1) return 0;
  const int p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - 1 - i;
}

// Whole-sample symmetric index (OpenCV BORDER_REFLECT_101), extended periodically.
inline int reflect101_index(int i, int n) {
  if (n == 1) return 0;
  const int p = 2 * n - 2;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - i;
}

// ---- Schmid bank and its symmetry classes ---------------------------------
// The 13 Schmid kernels are radially symmetric, so every tap with the same
// dx*dx+dy*dy carries the same weight.  A response is evaluated as
//   r_f = sum over classes c (ascending dx^2+dy^2) of  w_f[c] * S_c
// where S_c is the (exact, integer-valued) sum of grey values in class c, each
// term folded in with one fmaf.  This is the declared accumulation order that
// makes texton labels bit-reproducible between the GPU and the CPU.
struct SchmidTable {
  int num_classes = 0;
  std::vector<int> class_start;     // num_classes + 1
  std::vector<int> member_dx;       // 49*49 entries, grouped by class
  std::vector<int> member_dy;
  std::vector<float> weight;        // [f][class], kSchmidNF * num_classes
  std::vector<double> bank;         // [f][dy+24][dx+24], FP64 kernels
};
const SchmidTable& schmid_table();

// ---- orientation slices -----------------------------------------------------
// Slice index of the label at relative position (drow, dcol) from the centre,
// following orientation_slice_map (lib_image.cc:3052-3079) and the flipped
// window indexing of compute_hist_gradient_2D (lib_image.cc:3120-3231).
int slice_of_offset(int drow, int dcol);

// Histogram smoothing kernel: lib_image::gaussian(sigma) (lib_image.cc:1374-1436).
std::vector<double> gaussian_1d(double sigma);

// ---- nonmax suppression tables ------------------------------------------------
struct NonmaxTable {
  int branch[kNumOri];          // interpolation case per orientation index
  double d[kNumOri];            // interpolation weight per orientation index
  double cosdiff[kNumOri][kNumOri];  // cos of tolerance-corrected orientation difference [centre][neighbour]
};
const NonmaxTable& nonmax_table();

// ---- sPb oriented derivative filters -----------------------------------------
constexpr int kSPbRadius = 3;
constexpr int kSPbSize = 2 * kSPbRadius + 1;
// [o][drow+3][dcol+3], FP64, correlation taps
const std::vector<double>& spb_filters();

// ---- intervening-contour walk programs --------------------------------------
// For a pixel whose square window of radius 5 is clipped to the image at
// distances (left, right, top, bottom) in [0,5], and for each disc offset k
// (81 entries, row-major over dy then dx, disc dx^2+dy^2 <= 25), the program
// lists the pixel offsets whose mPb maximum is the intervening contour of that
// connection; `found` is false when buildW leaves the connection without a
// support entry (then the affinity uses maxpb = 1).
struct ICConnection {
  int8_t dx, dy;            // target offset
  uint8_t found;            // support entry present
  uint16_t site_start;      // into sites
  uint8_t site_count;
};
struct ICProgram {
  std::vector<ICConnection> conn;      // one per disc offset inside the clipped window
  std::vector<int8_t> site_dx, site_dy;
};
// Clip distance code: value in [0,5]; 5 means "not clipped".
const ICProgram& ic_program(int left, int right, int top, int bottom);
// The 81 disc offsets in row-major order (dy outer, dx inner).
void disc_offsets(int* dx, int* dy);  // 81 entries each
// The 40 "upper" offsets (dy > 0, or dy == 0 and dx > 0) in row-major order.
void upper_offsets(int* dx, int* dy);

// ---- synthetic images -------------------------------------------------------
struct SynthParams {
  int width = 321, height = 481;
  int regions = 60;
  double noise_sigma = 6.0;
  uint64_t seed = 1;
  int classes = 1;          // 1: uniform statistics; 3: three region classes
};
void synth_voronoi(const SynthParams& p, uint8_t* rgb /* H*W*3 */, int* region /* optional H*W */);
void make_codebook(const uint32_t* packed, int width, int height, uint64_t seed, int iters,
                   int max_samples, float* codebook /* 13 x 32 row-major */);

// FP32 Schmid responses at one pixel in the declared order (host mirror of the
// texton kernel; used for codebook training).
void schmid_response_at(const uint32_t* packed, int width, int height, int x, int y, float* r13);

}  // namespace gpb

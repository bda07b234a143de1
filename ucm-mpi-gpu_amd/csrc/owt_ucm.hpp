// OWT-UCM, the consumer of the 8 oriented gPb maps: the ultrametric contour
// map of src/contours2ucm.cpp (Contours2Ucm :405-490).  Host code: every stage
// after the 8-bit scaling is a sequential graph or geometry algorithm whose
// result depends on its visiting order (a level-by-level raster watershed,
// contour tracing and subdivision, greedy region merging), so it runs on the
// CPU in the reference's order, with data structures of its own.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace gpb {
namespace owt {

// Contours of a binary image (nonzero = contour pixel), as lib_image's
// contour_set builds them (globalPb/src/math/libraries/lib_image.cc:5320-5485
// vertices and traced edges, :5953-6044 subdivide_local with its default
// arguments, :6138-6326 subdivide_global).  Coordinates are (x, y) = (row,
// column); an edge's points are its interior pixels in traversal order.
struct Contours {
  std::vector<int32_t> vx, vy;                // vertex coordinates
  std::vector<int32_t> es, ee;                // edge start / end vertex ids
  std::vector<std::vector<int32_t>> px, py;   // interior points of each edge
  bool failed = false;                        // a subdivision could not be completed
};

// rows x cols row-major binary image -> contours after both subdivisions.
// When a subdivision cannot make the topology consistent the reference's
// caller catches the exception and continues with no contours
// (src/mex_contour_sides.cpp:216-223); then failed is set and the set is empty.
Contours fit_contours(const uint8_t* bw, int rows, int cols);

// The 8-connected watershed of src/watershed.cpp:28-108 (levels in increasing
// order, each level's pixels in raster order); labels > 0 for basins, 0 for
// watershed pixels.
std::vector<int32_t> watershed8(const uint8_t* a, int rows, int cols);

// Stages of Contours2Ucm on rows x cols maps (row-major, double).
struct UcmStages {
  std::vector<uint8_t> pb8;      // max over orientations scaled to 8 bits
  std::vector<int32_t> ws;       // watershed labels
  std::vector<double> ws_wt;     // contour strength of the finest partition
  std::vector<double> ws_wt2;    // (2 rows + 1) x (2 cols + 1): cleaned doubled grid
  std::vector<int32_t> labels;   // rows x cols initial regions
  std::vector<double> super_ucm; // (2 rows + 1) x (2 cols + 1), before calibration
  Contours contours;
  int regions = 0;
  // host milliseconds: 8-bit map, watershed, contours, strength, doubled grid
  // with cleaning and regions, merging, calibration
  double ms[7] = {0, 0, 0, 0, 0, 0, 0};
};

// The whole consumer: pb_oriented[8][rows][cols] -> ucm, (2 rows + 1) x
// (2 cols + 1) when double_size, else (rows + 1) x (cols + 1).  stages, if
// given, receives the intermediate results.
void contours2ucm(const double* pb_oriented, int rows, int cols, bool double_size, double* ucm,
                  UcmStages* stages = nullptr);

// The same consumer in three parts, so that the per-pixel parts can run
// where the maps are (the GPU path's device buffers):
//   1. pb8: the maximum over orientations scaled to 8 bits (pb8_from_maps on
//      the host);
//   2. finest_partition: watershed and contours of pb8 (host), and omask[N],
//      bit o set where a contour pixel belongs to an edge whose chord has
//      orientation o;
//   3. ws_wt[i] = max(0, pb_oriented[o][i] over the bits of omask[i])
//      (gather_strength on the host), then ucm_from_strength.
void pb8_from_maps(const double* pb_oriented, int rows, int cols, UcmStages& S);
void finest_partition(int rows, int cols, UcmStages& S, std::vector<uint8_t>& omask);
void gather_strength(const double* pb_oriented, const uint8_t* omask, size_t n, std::vector<double>& ws_wt);
void ucm_from_strength(int rows, int cols, bool double_size, double* ucm, UcmStages& S);

}  // namespace owt
}  // namespace gpb

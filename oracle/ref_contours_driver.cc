// Test-infrastructure driver around the reference's own contour extraction:
// lib_image::connected_components, lib_image::contour_set and its
// subdivide_local / subdivide_global, called the way MexContourSides calls
// them (src/mex_contour_sides.cpp:120-146, with completion_cdt left out: it
// only appends completion edges, which CreateFinestPartition skips,
// src/contours2ucm.cpp:339-341,368-369).  Compiled with lib_image from the
// reference tree by `make ref`; used only by tests, as the checker of the
// product's own contour code.
#include <cstdint>
#include <vector>

#include "lang/array.hh"
#include "lang/exceptions/exception.hh"
#include "math/libraries/lib_image.hh"
#include "math/matrices/matrix.hh"

using lang::array;
using math::libraries::lib_image;
using math::matrices::matrix;

namespace {
struct Result {
  std::vector<int32_t> vertices, edges, points;
  std::vector<int64_t> offsets;
  int failed = 0;
};
Result g_last;
}  // namespace

extern "C" {

// Runs the extraction on bw[H][W] (nonzero = contour) and keeps the result;
// returns the number of edges, or -1 where the reference throws (its caller
// then continues with no contours).  Read it back with ref_contours_get.
int ref_contours_run(const unsigned char* bw, int W, int H, int* n_vertices, long long* n_points) {
  g_last = Result();
  try {
    matrix<> im(static_cast<unsigned long>(H), static_cast<unsigned long>(W));
    for (int r = 0; r < H; r++)
      for (int c = 0; c < W; c++) im(unsigned(r), unsigned(c)) = bw[size_t(r) * W + c] ? 1.0 : 0.0;
    matrix<unsigned long> labels = lib_image::connected_components(im);
    lib_image::contour_set contours(labels);
    contours.subdivide_local();
    contours.subdivide_global();
    const unsigned long nv = contours.vertices_size(), ne = contours.edges_size();
    for (unsigned long v = 0; v < nv; v++) {
      g_last.vertices.push_back(int32_t(contours.vertex(v).x));
      g_last.vertices.push_back(int32_t(contours.vertex(v).y));
    }
    for (unsigned long e = 0; e < ne; e++) {
      const lib_image::contour_edge& ed = contours.edge(e);
      g_last.edges.push_back(int32_t(ed.vertex_start->id));
      g_last.edges.push_back(int32_t(ed.vertex_end->id));
      g_last.offsets.push_back(int64_t(g_last.points.size() / 2));
      for (unsigned long k = 0; k < ed.x_coords.size(); k++) {
        g_last.points.push_back(int32_t(ed.x_coords[k]));
        g_last.points.push_back(int32_t(ed.y_coords[k]));
      }
    }
    g_last.offsets.push_back(int64_t(g_last.points.size() / 2));
  } catch (lang::exceptions::exception&) {
    g_last = Result();
    g_last.failed = 1;
    *n_vertices = 0;
    *n_points = 0;
    return -1;
  }
  *n_vertices = int(g_last.vertices.size() / 2);
  *n_points = (long long)(g_last.points.size() / 2);
  return int(g_last.edges.size() / 2);
}

// Copies the last result: vertices[2 nv], edges[2 ne], offsets[ne + 1],
// points[2 np] (row, column pairs).
void ref_contours_get(int32_t* vertices, int32_t* edges, int64_t* offsets, int32_t* points) {
  for (size_t i = 0; i < g_last.vertices.size(); i++) vertices[i] = g_last.vertices[i];
  for (size_t i = 0; i < g_last.edges.size(); i++) edges[i] = g_last.edges[i];
  for (size_t i = 0; i < g_last.offsets.size(); i++) offsets[i] = g_last.offsets[i];
  for (size_t i = 0; i < g_last.points.size(); i++) points[i] = g_last.points[i];
}

}  // extern "C"

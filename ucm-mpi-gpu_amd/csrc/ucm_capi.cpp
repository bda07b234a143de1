// C entry points of the OWT-UCM consumer (owt_ucm.cpp): host buffers, no
// context (the reference's Contours2Ucm has none); each catches its own
// exceptions and returns a status.
#include <cstring>
#include <new>
#include <stdexcept>

#include "gpb_capi.h"
#include "owt_ucm.hpp"

namespace {

bool ucm_dims_ok(int width, int height) { return width >= 2 && height >= 2 && long(width) * height < (1L << 30); }

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return GPB_ERR_INTERNAL;
  } catch (const std::invalid_argument&) {
    return GPB_ERR_ARG;
  } catch (...) {
    return GPB_ERR_INTERNAL;
  }
}

}  // namespace

extern "C" {

int gpb_contours2ucm(int width, int height, const double* pb_oriented, int double_size, double* ucm) {
  if (!ucm_dims_ok(width, height) || !pb_oriented || !ucm) return GPB_ERR_ARG;
  return guarded([&] {
    gpb::owt::contours2ucm(pb_oriented, height, width, double_size != 0, ucm);
    return GPB_OK;
  });
}

int gpb_ucm_stages(int width, int height, const double* pb_oriented, uint8_t* pb8, int32_t* ws, double* ws_wt,
                   double* ws_wt2, int32_t* labels, double* super_ucm, int* regions, int* contours_failed,
                   double* stage_ms) {
  if (!ucm_dims_ok(width, height) || !pb_oriented) return GPB_ERR_ARG;
  return guarded([&] {
    gpb::owt::UcmStages st;
    const size_t N = size_t(width) * height, N2 = size_t(2 * width + 1) * (2 * height + 1);
    std::vector<double> u(N2);
    gpb::owt::contours2ucm(pb_oriented, height, width, true, u.data(), &st);
    if (pb8) std::memcpy(pb8, st.pb8.data(), N);
    if (ws) std::memcpy(ws, st.ws.data(), N * sizeof(int32_t));
    if (ws_wt) std::memcpy(ws_wt, st.ws_wt.data(), N * sizeof(double));
    if (ws_wt2) std::memcpy(ws_wt2, st.ws_wt2.data(), N2 * sizeof(double));
    if (labels) std::memcpy(labels, st.labels.data(), N * sizeof(int32_t));
    if (super_ucm) std::memcpy(super_ucm, st.super_ucm.data(), N2 * sizeof(double));
    if (regions) *regions = st.regions;
    if (contours_failed) *contours_failed = st.contours.failed ? 1 : 0;
    if (stage_ms)
      for (int k = 0; k < 7; k++) stage_ms[k] = st.ms[k];
    return GPB_OK;
  });
}

int gpb_watershed8(int width, int height, const uint8_t* image, int32_t* labels) {
  if (!ucm_dims_ok(width, height) || !image || !labels) return GPB_ERR_ARG;
  return guarded([&] {
    const std::vector<int32_t> L = gpb::owt::watershed8(image, height, width);
    std::memcpy(labels, L.data(), L.size() * sizeof(int32_t));
    return GPB_OK;
  });
}

int gpb_fit_contours(int width, int height, const uint8_t* bw, int max_vertices, int32_t* vertices, int max_edges,
                     int32_t* edges, int64_t* point_offsets, int64_t max_points, int32_t* points, int* n_vertices,
                     int* n_edges, int64_t* n_points) {
  if (!ucm_dims_ok(width, height) || !bw || !n_vertices || !n_edges || !n_points) return GPB_ERR_ARG;
  return guarded([&] {
    const gpb::owt::Contours c = gpb::owt::fit_contours(bw, height, width);
    const int nv = int(c.vx.size()), ne = int(c.es.size());
    int64_t np = 0;
    for (const auto& p : c.px) np += int64_t(p.size());
    *n_vertices = nv;
    *n_edges = ne;
    *n_points = np;
    if (c.failed) return GPB_ERR_INTERNAL;
    if (nv > max_vertices || ne > max_edges || np > max_points || (nv && !vertices) || (ne && (!edges || !point_offsets)) ||
        (np && !points))
      return GPB_ERR_TOO_LARGE;
    for (int v = 0; v < nv; v++) {
      vertices[2 * v] = c.vx[size_t(v)];
      vertices[2 * v + 1] = c.vy[size_t(v)];
    }
    int64_t at = 0;
    for (int e = 0; e < ne; e++) {
      edges[2 * e] = c.es[size_t(e)];
      edges[2 * e + 1] = c.ee[size_t(e)];
      point_offsets[e] = at;
      for (size_t k = 0; k < c.px[size_t(e)].size(); k++, at++) {
        points[2 * at] = c.px[size_t(e)][k];
        points[2 * at + 1] = c.py[size_t(e)][k];
      }
    }
    if (ne) point_offsets[ne] = at;
    return GPB_OK;
  });
}

}  // extern "C"

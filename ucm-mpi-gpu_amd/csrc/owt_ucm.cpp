// OWT-UCM on the host: finest partition, doubled grid, greedy merging by mean
// boundary strength.  The result must equal the reference's bit for bit, so
// every floating-point expression keeps the reference's operand order (this
// file is built with -ffp-contract=off) and every tie is broken the way the
// reference's visiting order breaks it; the data structures are this file's
// own (counting-sorted watershed, union-find components, flat edge arrays,
// an exact integer crossing search).
#include "owt_ucm.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <unordered_map>

namespace gpb {
namespace owt {
namespace {

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------------------
// Contour geometry
// ---------------------------------------------------------------------------

// Distance from (x, y) to the segment p-q, with the operations of
// point_2D::distance_to_segment (globalPb/src/math/geometry/point_2D.cc:239-257).
double dist_to_segment(double x, double y, double px, double py, double qx, double qy) {
  const double xa = qx - px, ya = qy - py;
  const double xb = x - px, yb = y - py;
  const double xc = x - qx, yc = y - qy;
  if ((xa * xc + ya * yc) >= 0) return std::sqrt(xc * xc + yc * yc);
  if ((-xa * xb - ya * yb) >= 0) return std::sqrt(xb * xb + yb * yb);
  return std::fabs(xa * yb - ya * xb) / std::sqrt(xa * xa + ya * ya);
}

// Whether a contour pixel must be a vertex given which of its 8 neighbours
// are contour pixels (n/s: column +-1, e/w: row +-1), the conditions of
// lib_image.cc:4421-4478: ends (at most one neighbour, or two that touch),
// 2x2 blocks, junctions, and everything but the middle of an H.
bool must_be_vertex(bool n, bool s, bool e, bool w, bool ne, bool nw, bool se, bool sw) {
  const int adj = n + s + e + w, diag = ne + nw + se + sw, all = adj + diag;
  if (all <= 1) return true;
  if (all == 2) return (n && (ne || nw)) || (s && (se || sw)) || (e && (ne || se)) || (w && (nw || sw));
  if (all <= 4) {
    if ((nw && n && w) || (n && ne && e) || (w && sw && s) || (e && s && se)) return true;
    if ((nw && n && ne) || (sw && s && se) || (nw && w && sw) || (ne && e && se)) return all == 3;
    if (adj >= 3 || diag >= 3) return true;
    if ((n && se && sw) || (s && ne && nw) || (w && se && ne) || (e && nw && sw)) return true;
    return (nw && s && e) || (ne && s && w) || (sw && n && e) || (se && n && w);
  }
  return !(adj == 2 && ((n && s) || (e && w)));
}

struct Builder {
  int sx, sy;  // rows, columns
  const uint8_t* bw;
  std::vector<uint8_t> is_vertex, is_edge;
  std::vector<int32_t> assign;
  Contours c;

  bool on(long x, long y) const { return bw[size_t(x) * sy + y] != 0; }

  int add_vertex(long x, long y) {
    c.vx.push_back(int32_t(x));
    c.vy.push_back(int32_t(y));
    return int(c.vx.size()) - 1;
  }
  int add_edge(int a, int b) {
    c.es.push_back(a);
    c.ee.push_back(b);
    c.px.emplace_back();
    c.py.emplace_back();
    return int(c.es.size()) - 1;
  }

  // Trace the contour that leaves vertex v through pixel (x, y), preferring
  // an adjacent pixel that is not a vertex, then an adjacent vertex, then the
  // same for diagonal pixels, each in the order n, s, e, w (ne, nw, se, sw);
  // where none is free the last pixel becomes a new vertex (the steps of
  // contour_set::trace_edge, lib_image.cc:5550-5680).
  void trace(int v, long x, long y) {
    const size_t iv = size_t(c.vx[v]) * sy + c.vy[v];
    is_edge[iv] = 1;  // the start vertex is not a step target while tracing
    const int e_id = int(c.es.size());
    std::vector<int32_t> tx, ty;
    bool new_vertex = false;
    while (true) {
      tx.push_back(int32_t(x));
      ty.push_back(int32_t(y));
      const size_t i = size_t(x) * sy + y;
      is_edge[i] = 1;
      assign[i] = e_id;
      const bool hn = y + 1 < sy, hs = y > 0, he = x + 1 < sx, hw = x > 0;
      auto free = [&](bool has, size_t j) { return has && bw[j] != 0 && !is_edge[j]; };
      const size_t jn = i + 1, js = i - 1, je = i + sy, jw = i - sy;
      const bool an = free(hn, jn), as = free(hs, js), ae = free(he, je), aw = free(hw, jw);
      const bool ane = free(hn && he, je + 1), anw = free(hn && hw, jw + 1);
      const bool ase = free(hs && he, je - 1), asw = free(hs && hw, jw - 1);
      // adjacent pixel that is not a vertex: keep tracing
      if (an && !is_vertex[jn]) { y++; continue; }
      if (as && !is_vertex[js]) { y--; continue; }
      if (ae && !is_vertex[je]) { x++; continue; }
      if (aw && !is_vertex[jw]) { x--; continue; }
      // adjacent vertex: the edge ends there
      if (an) { y++; break; }
      if (as) { y--; break; }
      if (ae) { x++; break; }
      if (aw) { x--; break; }
      if (ane && !is_vertex[je + 1]) { x++; y++; continue; }
      if (anw && !is_vertex[jw + 1]) { x--; y++; continue; }
      if (ase && !is_vertex[je - 1]) { x++; y--; continue; }
      if (asw && !is_vertex[jw - 1]) { x--; y--; continue; }
      if (ane) { x++; y++; break; }
      if (anw) { x--; y++; break; }
      if (ase) { x++; y--; break; }
      if (asw) { x--; y--; break; }
      new_vertex = true;  // a dead end: this pixel is the end vertex
      is_edge[i] = 0;
      break;
    }
    is_edge[iv] = 0;
    const size_t iend = size_t(x) * sy + y;
    if (new_vertex) {
      const int nv = add_vertex(x, y);
      assign[iend] = nv;
      is_vertex[iend] = 1;
    }
    const int e = add_edge(v, assign[iend]);
    const size_t n = new_vertex ? tx.size() - 1 : tx.size();
    c.px[e].assign(tx.begin(), tx.begin() + long(n));
    c.py[e].assign(ty.begin(), ty.begin() + long(n));
  }

  void build() {
    const size_t N = size_t(sx) * sy;
    is_vertex.assign(N, 0);
    is_edge.assign(N, 0);
    assign.assign(N, 0);
    // 8-connected components, numbered by their first pixel in raster order
    // (lib_image::connected_components, lib_image.cc:4773-4900); a component
    // without a vertex (a closed loop) gets one at its last pixel.
    std::vector<int32_t> parent(N, -1);
    auto find = [&](int32_t a) {
      while (parent[size_t(a)] != a) {
        parent[size_t(a)] = parent[size_t(parent[size_t(a)])];
        a = parent[size_t(a)];
      }
      return a;
    };
    for (long x = 0; x < sx; x++)
      for (long y = 0; y < sy; y++) {
        const size_t i = size_t(x) * sy + y;
        if (!bw[i]) continue;
        parent[i] = int32_t(i);
        const long nb[4][2] = {{x - 1, y - 1}, {x - 1, y}, {x - 1, y + 1}, {x, y - 1}};
        for (const auto& q : nb) {
          if (q[0] < 0 || q[1] < 0 || q[1] >= sy || !on(q[0], q[1])) continue;
          const int32_t a = find(int32_t(i)), b = find(int32_t(size_t(q[0]) * sy + q[1]));
          if (a != b) parent[size_t(std::max(a, b))] = std::min(a, b);
        }
      }
    // unions keep the smaller root: a component's root is its first pixel
    std::vector<int32_t> comp(N, -1);
    std::vector<size_t> last;
    std::vector<uint8_t> has_vertex;
    for (long x = 0; x < sx; x++)
      for (long y = 0; y < sy; y++) {
        const size_t i = size_t(x) * sy + y;
        if (!bw[i]) continue;
        const int32_t r = find(int32_t(i));
        int32_t k;
        if (size_t(r) == i) {
          k = int32_t(last.size());
          comp[i] = k;
          last.push_back(i);
          has_vertex.push_back(0);
        } else {
          k = comp[size_t(r)];
          last[size_t(k)] = i;
        }
        const bool hn = y + 1 < sy, hs = y > 0, he = x + 1 < sx, hw = x > 0;
        const bool n = hn && on(x, y + 1), s = hs && on(x, y - 1), e = he && on(x + 1, y), w = hw && on(x - 1, y);
        const bool ne = hn && he && on(x + 1, y + 1), nw = hn && hw && on(x - 1, y + 1);
        const bool se = hs && he && on(x + 1, y - 1), sw = hs && hw && on(x - 1, y - 1);
        if (must_be_vertex(n, s, e, w, ne, nw, se, sw)) {
          assign[i] = add_vertex(x, y);
          is_vertex[i] = 1;
          has_vertex[size_t(k)] = 1;
        }
      }
    for (size_t k = 0; k < last.size(); k++)
      if (!has_vertex[k]) {
        const size_t i = last[k];
        assign[i] = add_vertex(long(i / size_t(sy)), long(i % size_t(sy)));
        is_vertex[i] = 1;
      }
    // from each vertex (new end vertices join the list as they appear):
    // adjacent contour pixels, then diagonal ones not reachable through them
    for (size_t v = 0; v < c.vx.size(); v++) {
      const long X = c.vx[v], Y = c.vy[v];
      long qx[8], qy[8];
      int nq = 0;
      const long xs = X > 0 ? X - 1 : X + 1, xe = X + 1;
      const long ys = Y > 0 ? Y - 1 : Y + 1, ye = Y + 1;
      for (long x = xs; x <= xe && x < sx; x += 2)
        if (on(x, Y)) { qx[nq] = x; qy[nq++] = Y; }
      for (long y = ys; y <= ye && y < sy; y += 2)
        if (on(X, y)) { qx[nq] = X; qy[nq++] = y; }
      for (long x = xs; x <= xe && x < sx; x += 2)
        for (long y = ys; y <= ye && y < sy; y += 2)
          if (on(x, y) && !on(X, y) && !on(x, Y)) { qx[nq] = x; qy[nq++] = y; }
      for (int k = 0; k < nq; k++) {
        const size_t i = size_t(qx[k]) * sy + qy[k];
        if (is_vertex[i]) {
          if (size_t(assign[i]) > v) add_edge(int(v), assign[i]);
        } else if (!is_edge[i]) {
          trace(int(v), qx[k], qy[k]);
        }
      }
    }
    // drop vertices without edges (single-pixel islands), keeping the order
    std::vector<int32_t> deg(c.vx.size(), 0);
    for (size_t e = 0; e < c.es.size(); e++) {
      deg[size_t(c.es[e])]++;
      deg[size_t(c.ee[e])]++;
    }
    std::vector<int32_t> remap(c.vx.size(), -1);
    size_t nv = 0;
    for (size_t v = 0; v < c.vx.size(); v++)
      if (deg[v]) {
        remap[v] = int32_t(nv);
        c.vx[nv] = c.vx[v];
        c.vy[nv] = c.vy[v];
        nv++;
      }
    c.vx.resize(nv);
    c.vy.resize(nv);
    for (size_t e = 0; e < c.es.size(); e++) {
      c.es[e] = remap[size_t(c.es[e])];
      c.ee[e] = remap[size_t(c.ee[e])];
    }
  }

  // Split edge e at its interior point p: a new vertex there, e keeps the
  // points before it and a new edge (to e's old end) the points after
  // (contour_set::split_edge, lib_image.cc:5685-5740).
  void split(int e, size_t p) {
    const int v = add_vertex(c.px[size_t(e)][p], c.py[size_t(e)][p]);
    const int old_end = c.ee[size_t(e)];
    const int ne = add_edge(v, old_end);
    c.ee[size_t(e)] = v;
    auto& ox = c.px[size_t(e)];
    auto& oy = c.py[size_t(e)];
    c.px[size_t(ne)].assign(ox.begin() + long(p) + 1, ox.end());
    c.py[size_t(ne)].assign(oy.begin() + long(p) + 1, oy.end());
    ox.resize(p);
    oy.resize(p);
  }

  double sxv(int v) const { return double(c.vx[size_t(v)]); }
  double syv(int v) const { return double(c.vy[size_t(v)]); }

  // Interior point of e farthest from the segment p-q (the first of equals).
  size_t farthest(int e, double px, double py, double qx, double qy) const {
    const auto& ex = c.px[size_t(e)];
    const auto& ey = c.py[size_t(e)];
    size_t pos = 0;
    double best = 0;
    for (size_t n = 0; n < ex.size(); n++) {
      const double d = dist_to_segment(ex[n], ey[n], px, py, qx, qy);
      if (d > best) {
        pos = n;
        best = d;
      }
    }
    return pos;
  }

  // Interior point of e closest to the segment p-q (the first of equals).
  size_t closest(int e, double px, double py, double qx, double qy) const {
    const auto& ex = c.px[size_t(e)];
    const auto& ey = c.py[size_t(e)];
    size_t pos = 0;
    double best = dist_to_segment(ex[0], ey[0], px, py, qx, qy);
    for (size_t n = 1; n < ex.size(); n++) {
      const double d = dist_to_segment(ex[n], ey[n], px, py, qx, qy);
      if (d < best) {
        pos = n;
        best = d;
      }
    }
    return pos;
  }

  // Split until each edge lies close to its chord: a point farther than 5%
  // of the chord and 2 pixels from it, or seen from either end at more than
  // 90 degrees from the chord, splits the edge at its farthest point
  // (contour_set::subdivide_local with its defaults, lib_image.cc:5953-6044,
  // lib_image.hh:1521-1527).  Each edge's subdivision depends only on its own
  // points, and a split edge and its new half go to the back of the queue.
  void subdivide_local() {
    const double max_dist_dev = 0.05, tol_dist = 2.0, max_angle = M_PI_2;
    std::deque<int> q;
    for (size_t e = 0; e < c.es.size(); e++) q.push_back(int(e));
    while (!q.empty()) {
      const int e = q.front();
      q.pop_front();
      const double x0 = sxv(c.es[size_t(e)]), y0 = syv(c.es[size_t(e)]);
      const double x1 = sxv(c.ee[size_t(e)]), y1 = syv(c.ee[size_t(e)]);
      const double rx = x1 - x0, ry = y1 - y0;
      const double len = std::sqrt(rx * rx + ry * ry);
      const double max_dist = max_dist_dev * len;
      size_t pos = 0;
      double best = 0;
      bool do_split = false;
      const auto& ex = c.px[size_t(e)];
      const auto& ey = c.py[size_t(e)];
      for (size_t n = 0; n < ex.size(); n++) {
        const double x = ex[n], y = ey[n];
        const double d = dist_to_segment(x, y, x0, y0, x1, y1);
        if (d > best) {
          pos = n;
          best = d;
        }
        if (!do_split) do_split = (d > max_dist) && (d > tol_dist);
        if (!do_split) {
          const double ax = x - x0, ay = y - y0, bx = x - x1, by = y - y1;
          const double a0 = std::acos((ax * rx + ay * ry) / (std::sqrt(ax * ax + ay * ay) * len));
          const double a1 = std::acos((bx * -rx + by * -ry) / (std::sqrt(bx * bx + by * by) * len));
          do_split = ((a0 > a1) ? a0 : a1) > max_angle;
        }
      }
      if (do_split) {
        split(e, pos);
        q.push_back(e);
        q.push_back(int(c.es.size()) - 1);
      }
    }
  }

  // Order of two edges by their (smaller, larger) vertex ids.
  int key_cmp(int a, int b) const {
    const int a0 = std::min(c.es[size_t(a)], c.ee[size_t(a)]), a1 = std::max(c.es[size_t(a)], c.ee[size_t(a)]);
    const int b0 = std::min(c.es[size_t(b)], c.ee[size_t(b)]), b1 = std::max(c.es[size_t(b)], c.ee[size_t(b)]);
    if (a0 != b0) return a0 < b0 ? -1 : 1;
    return a1 < b1 ? -1 : (a1 > b1 ? 1 : 0);
  }

  // The library's array sort (globalPb/include/lang/array.hh:1676-1710): a
  // quicksort on the middle element whose index follows it through swaps.
  // Equal keys (edges joining the same two vertices) keep the relative order
  // this partition leaves, which decides which of them is split.
  void ref_sort(std::vector<int>& d, long lo, long hi) const {
    if (lo >= hi) return;
    long pivot = (lo + hi) / 2, i = lo, j = hi;
    while (true) {
      while (key_cmp(d[size_t(i)], d[size_t(pivot)]) < 0) i++;
      while (key_cmp(d[size_t(j)], d[size_t(pivot)]) > 0) j--;
      if (i < j) {
        std::swap(d[size_t(i)], d[size_t(j)]);
        if (pivot == i)
          pivot = j;
        else if (pivot == j)
          pivot = i;
        i++;
        j--;
      } else {
        break;
      }
    }
    ref_sort(d, lo, j);
    ref_sort(d, j + 1, hi);
  }

  struct Crossing {
    std::vector<int> inter;  // edges whose chord passes through the point's interior
    bool at_vertex = false;  // the point is a vertex (some chord ends there)
    double x = 0, y = 0;
  };

  // Make the chords' topology that of the contours (contour_set::
  // subdivide_global, lib_image.cc:6138-6326): first split one of any two
  // edges joining the same vertices (the less straight one), then, until no
  // chord passes through another's interior, split the chords that do: at a
  // vertex every chord through it, at a crossing of chords all but one (for
  // two, the one whose contour lies closer to the other's chord is kept).
  // Returns false where the reference throws.
  bool subdivide_global() {
    {
      std::vector<int> sorted(c.es.size());
      std::iota(sorted.begin(), sorted.end(), 0);
      if (sorted.size() > 1) ref_sort(sorted, 0, long(sorted.size()) - 1);
      size_t prev = 0;
      for (size_t cur = 1; cur < sorted.size(); cur++) {
        const int ep = sorted[prev], ec = sorted[cur];
        if (key_cmp(ep, ec) != 0) {
          prev = cur;
          continue;
        }
        const double px = sxv(c.es[size_t(ep)]), py = syv(c.es[size_t(ep)]);
        const double qx = sxv(c.ee[size_t(ep)]), qy = syv(c.ee[size_t(ep)]);
        const bool can_p = !c.px[size_t(ep)].empty(), can_c = !c.px[size_t(ec)].empty();
        bool split_prev = false;
        if (can_p && can_c) {
          const size_t ip = farthest(ep, px, py, qx, qy), ic = farthest(ec, px, py, qx, qy);
          const double dp = dist_to_segment(c.px[size_t(ep)][ip], c.py[size_t(ep)][ip], px, py, qx, qy);
          const double dc = dist_to_segment(c.px[size_t(ec)][ic], c.py[size_t(ec)][ic], px, py, qx, qy);
          split_prev = dp > dc;
        } else if (can_p) {
          split_prev = true;
        } else if (!can_c) {
          return false;
        }
        const int e = split_prev ? ep : ec;
        split(e, farthest(e, px, py, qx, qy));
        if (split_prev) prev = cur;
      }
    }
    while (true) {
      std::vector<Crossing> xs;
      find_crossings(xs);
      if (xs.empty()) break;
      const size_t n_edges = c.es.size();
      std::vector<uint8_t> mark(n_edges, 0);
      for (auto& cr : xs) {
        if (cr.at_vertex) {
          for (int e : cr.inter) mark[size_t(e)] = 1;
          continue;
        }
        int ea = cr.inter[0];
        for (size_t k = 1; k < cr.inter.size(); k++) {
          const int eb = cr.inter[k];
          if (c.px[size_t(ea)].empty()) {
            mark[size_t(eb)] = 1;
            continue;
          }
          if (c.px[size_t(eb)].empty()) {
            mark[size_t(ea)] = 1;
            ea = eb;
            continue;
          }
          const double pax = sxv(c.es[size_t(ea)]), pay = syv(c.es[size_t(ea)]);
          const double qax = sxv(c.ee[size_t(ea)]), qay = syv(c.ee[size_t(ea)]);
          const double pbx = sxv(c.es[size_t(eb)]), pby = syv(c.es[size_t(eb)]);
          const double qbx = sxv(c.ee[size_t(eb)]), qby = syv(c.ee[size_t(eb)]);
          const size_t ia = closest(ea, pbx, pby, qbx, qby);
          const double da = dist_to_segment(c.px[size_t(ea)][ia], c.py[size_t(ea)][ia], pbx, pby, qbx, qby);
          const size_t ib = closest(eb, pax, pay, qax, qay);
          const double db = dist_to_segment(c.px[size_t(eb)][ib], c.py[size_t(eb)][ib], pax, pay, qax, qay);
          if (da < db) {
            mark[size_t(eb)] = 1;
          } else {
            mark[size_t(ea)] = 1;
            ea = eb;
          }
        }
      }
      for (size_t e = 0; e < n_edges; e++) {
        if (!mark[e]) continue;
        if (c.px[e].empty()) return false;
        const double px = sxv(c.es[e]), py = syv(c.es[e]), qx = sxv(c.ee[e]), qy = syv(c.ee[e]);
        split(int(e), farthest(int(e), px, py, qx, qy));
      }
    }
    return true;
  }

  // Every point through which some chord passes in its interior, found with
  // exact integer orientation tests over the chords sharing a grid cell:
  // vertices lying inside a chord, and crossings of two chords' interiors
  // (grouped by their exact rational position).  The chords through a
  // crossing are ordered as the reference's sweep orders them there (by the
  // angle of the part right of the sweep line, then by id).
  void find_crossings(std::vector<Crossing>& out) const {
    const size_t ne = c.es.size();
    constexpr int kCell = 8;
    const int gx = sx / kCell + 1, gy = sy / kCell + 1;
    std::vector<std::vector<int>> cells(size_t(gx) * gy);
    for (size_t e = 0; e < ne; e++) {
      const int a = c.es[e], b = c.ee[e];
      const int x0 = std::min(c.vx[size_t(a)], c.vx[size_t(b)]) / kCell, x1 = std::max(c.vx[size_t(a)], c.vx[size_t(b)]) / kCell;
      const int y0 = std::min(c.vy[size_t(a)], c.vy[size_t(b)]) / kCell, y1 = std::max(c.vy[size_t(a)], c.vy[size_t(b)]) / kCell;
      for (int x = x0; x <= x1; x++)
        for (int y = y0; y <= y1; y++) cells[size_t(x) * gy + y].push_back(int(e));
    }
    std::vector<uint64_t> pairs;
    for (const auto& cell : cells)
      for (size_t i = 0; i < cell.size(); i++)
        for (size_t j = i + 1; j < cell.size(); j++) {
          const uint64_t a = uint64_t(std::min(cell[i], cell[j])), b = uint64_t(std::max(cell[i], cell[j]));
          pairs.push_back((a << 32) | b);
        }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    struct Key {
      int64_t xn, yn, d;
      bool operator<(const Key& o) const {
        // lexicographic on the exact point (xn/d, yn/d), d > 0
        const __int128 lx = __int128(xn) * o.d, rx = __int128(o.xn) * d;
        if (lx != rx) return lx < rx;
        return __int128(yn) * o.d < __int128(o.yn) * d;
      }
    };
    std::map<Key, size_t> index;
    auto at = [&](const Key& k, bool vertex) -> Crossing& {
      auto it = index.find(k);
      if (it == index.end()) {
        it = index.emplace(k, out.size()).first;
        out.emplace_back();
        out.back().x = double(k.xn) / double(k.d);
        out.back().y = double(k.yn) / double(k.d);
      }
      Crossing& cr = out[it->second];
      cr.at_vertex = cr.at_vertex || vertex;
      return cr;
    };
    auto orient = [](int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t px, int64_t py) {
      const int64_t v = (bx - ax) * (py - ay) - (by - ay) * (px - ax);
      return (v > 0) - (v < 0);
    };
    auto strictly_inside = [](int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t px, int64_t py) {
      // p on the line through a, b (checked by the caller) and strictly between them
      if ((px == ax && py == ay) || (px == bx && py == by)) return false;
      return std::min(ax, bx) <= px && px <= std::max(ax, bx) && std::min(ay, by) <= py && py <= std::max(ay, by);
    };
    auto add_inter = [](Crossing& cr, int e) {
      if (std::find(cr.inter.begin(), cr.inter.end(), e) == cr.inter.end()) cr.inter.push_back(e);
    };
    for (uint64_t pr : pairs) {
      const int e0 = int(pr >> 32), e1 = int(pr & 0xffffffffu);
      const int64_t ax = c.vx[size_t(c.es[size_t(e0)])], ay = c.vy[size_t(c.es[size_t(e0)])];
      const int64_t bx = c.vx[size_t(c.ee[size_t(e0)])], by = c.vy[size_t(c.ee[size_t(e0)])];
      const int64_t cx = c.vx[size_t(c.es[size_t(e1)])], cy = c.vy[size_t(c.es[size_t(e1)])];
      const int64_t dx = c.vx[size_t(c.ee[size_t(e1)])], dy = c.vy[size_t(c.ee[size_t(e1)])];
      const int o_abc = orient(ax, ay, bx, by, cx, cy), o_abd = orient(ax, ay, bx, by, dx, dy);
      const int o_cda = orient(cx, cy, dx, dy, ax, ay), o_cdb = orient(cx, cy, dx, dy, bx, by);
      // an endpoint of one chord inside the other
      if (o_abc == 0 && strictly_inside(ax, ay, bx, by, cx, cy)) add_inter(at({cx, cy, 1}, true), e0);
      if (o_abd == 0 && strictly_inside(ax, ay, bx, by, dx, dy)) add_inter(at({dx, dy, 1}, true), e0);
      if (o_cda == 0 && strictly_inside(cx, cy, dx, dy, ax, ay)) add_inter(at({ax, ay, 1}, true), e1);
      if (o_cdb == 0 && strictly_inside(cx, cy, dx, dy, bx, by)) add_inter(at({bx, by, 1}, true), e1);
      // interiors crossing
      if (o_abc * o_abd < 0 && o_cda * o_cdb < 0) {
        // a + t (b - a), t = num / den
        const int64_t rx = bx - ax, ry = by - ay, sx_ = dx - cx, sy_ = dy - cy;
        int64_t den = rx * sy_ - ry * sx_;
        int64_t num = (cx - ax) * sy_ - (cy - ay) * sx_;
        if (den < 0) {
          den = -den;
          num = -num;
        }
        int64_t xn = ax * den + num * rx, yn = ay * den + num * ry, d = den;
        const int64_t g = std::gcd(std::gcd(xn < 0 ? -xn : xn, yn < 0 ? -yn : yn), d);
        if (g > 1) {
          xn /= g;
          yn /= g;
          d /= g;
        }
        const bool vertex_pt = false;  // a vertex inside both chords is found above
        if (d == 1) {
          // an integer point: a vertex there makes it a vertex crossing
          Crossing& cr = at({xn, yn, 1}, vertex_pt);
          add_inter(cr, e0);
          add_inter(cr, e1);
        } else {
          Crossing& cr = at({xn, yn, d}, false);
          add_inter(cr, e0);
          add_inter(cr, e1);
        }
      }
    }
    // an integer crossing point that is a vertex: chords end there
    std::unordered_map<int64_t, int> vat;
    for (size_t v = 0; v < c.vx.size(); v++) vat.emplace(int64_t(c.vx[v]) * (int64_t(sy) + 1) + c.vy[v], int(v));
    for (auto& cr : out) {
      if (cr.at_vertex) continue;
      if (cr.x == std::floor(cr.x) && cr.y == std::floor(cr.y) &&
          vat.count(int64_t(cr.x) * (int64_t(sy) + 1) + int64_t(cr.y)))
        cr.at_vertex = true;
    }
    // sweep order of the chords through a crossing
    for (auto& cr : out) {
      if (cr.at_vertex || cr.inter.size() < 2) continue;
      auto ang = [&](int e) {
        const int a = c.es[size_t(e)], b = c.ee[size_t(e)];
        const bool a_lower = (c.vx[size_t(a)] < c.vx[size_t(b)]) ||
                             (c.vx[size_t(a)] == c.vx[size_t(b)] && c.vy[size_t(a)] < c.vy[size_t(b)]);
        const int up = a_lower ? b : a;
        return std::atan2(double(c.vx[size_t(up)]) - cr.x, cr.y - double(c.vy[size_t(up)]));
      };
      std::vector<std::pair<double, int>> keyed;
      for (int e : cr.inter) keyed.emplace_back(ang(e), e);
      std::sort(keyed.begin(), keyed.end());
      for (size_t k = 0; k < keyed.size(); k++) cr.inter[k] = keyed[k].second;
    }
    // the sweep visits crossings in lexicographic order of their position
    std::vector<size_t> order(out.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
      return out[a].x < out[b].x || (out[a].x == out[b].x && out[a].y < out[b].y);
    });
    std::vector<Crossing> sorted;
    sorted.reserve(out.size());
    for (size_t k : order)
      if (!out[k].inter.empty()) sorted.push_back(std::move(out[k]));
    out.swap(sorted);
  }
};

// ---------------------------------------------------------------------------
// Regions
// ---------------------------------------------------------------------------

// Region ids of the pixels from the cleaned doubled grid g (R2 x C2): the
// 8-connected components of its zero sites, numbered from 0 in the order of
// their first site in raster order (bwlabel's numbering, less one), read at
// the odd sites (2i + 1, 2j + 1).  Components are built from runs of zeros:
// two runs in adjacent rows touch when their column ranges, one widened by a
// site on each side, overlap.  labels[i][j] = -1 where the site is not zero.
// Returns the number of components.
int region_labels(const double* g, int R2, int C2, int rows, int cols, int32_t* labels) {
  std::vector<std::vector<std::pair<int32_t, int32_t>>> row_runs(static_cast<size_t>(R2));
#pragma omp parallel for schedule(static)
  for (int r = 0; r < R2; r++) {
    const double* gr = g + size_t(r) * C2;
    auto& rr = row_runs[size_t(r)];
    for (int c = 0; c < C2;) {
      if (gr[c] != 0) {
        c++;
        continue;
      }
      const int s0 = c;
      while (c < C2 && gr[c] == 0) c++;
      rr.emplace_back(s0, c - 1);
    }
  }
  std::vector<int32_t> first(static_cast<size_t>(R2) + 1, 0);
  for (int r = 0; r < R2; r++) first[size_t(r) + 1] = first[size_t(r)] + int32_t(row_runs[size_t(r)].size());
  const int32_t nrun = first[size_t(R2)];
  std::vector<int32_t> parent(static_cast<size_t>(nrun));
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int32_t a) {
    while (parent[size_t(a)] != a) {
      parent[size_t(a)] = parent[size_t(parent[size_t(a)])];
      a = parent[size_t(a)];
    }
    return a;
  };
  for (int r = 1; r < R2; r++) {
    const auto& up = row_runs[size_t(r - 1)];
    const auto& cur = row_runs[size_t(r)];
    size_t u = 0;
    for (size_t k = 0; k < cur.size(); k++) {
      const int32_t s0 = cur[k].first, e0 = cur[k].second;
      while (u < up.size() && up[u].second + 1 < s0) u++;
      for (size_t v = u; v < up.size() && up[v].first <= e0 + 1; v++) {
        const int32_t a = find(first[size_t(r) - 1] + int32_t(v)), b = find(first[size_t(r)] + int32_t(k));
        if (a != b) parent[size_t(std::max(a, b))] = std::min(a, b);
      }
    }
  }
  // unions keep the smaller root: a component's root is its first run
  std::vector<int32_t> lab(static_cast<size_t>(nrun));
  int32_t next = 0;
  for (int32_t k = 0; k < nrun; k++) {
    const int32_t rt = find(k);
    lab[size_t(k)] = (rt == k) ? next++ : lab[size_t(rt)];
  }
  std::fill(labels, labels + size_t(rows) * cols, -1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < rows; i++) {
    const int r = 2 * i + 1;
    const auto& rr = row_runs[size_t(r)];
    for (size_t k = 0; k < rr.size(); k++) {
      const int32_t l = lab[size_t(first[size_t(r)]) + k];
      for (int c = rr[k].first | 1; c <= rr[k].second; c += 2)
        if ((c - 1) / 2 < cols) labels[size_t(i) * cols + (c - 1) / 2] = l;
    }
  }
  return next;
}

// Greedy merging of the initial regions by mean boundary strength
// (ComputeUcm, src/ucm_mean_pb.cpp:254-399).  The pair of adjacent regions
// whose boundary has the smallest mean strength merges first (ties: the queue
// entry with the smaller first, then second, region id); the merged region
// keeps the larger id, and its sums of strength and length with each
// neighbour are the sums of the two it came from.  After every merge the
// reference pushes a queue entry for every neighbour of the merged region,
// changed or not, and the region order inside an entry breaks ties; this
// pushes the same entries except those an identical or earlier-ordered live
// entry of the same pair already shadows, so the merge order is the
// reference's, ties included.
//
// The reference stores each region's boundary elements and, at every merge,
// raises those of the absorbed region to the merge's strength.  Merge
// strengths never decrease (a merged boundary's mean lies between the two
// means it combines, both at least the current minimum; the reference treats
// a decrease as an error), so an element ends at the strength of the merge
// that joins its two sides.  Here each pair of adjacent regions keeps the
// list of elements between them, and those take the strength when the pair
// merges.  Labels are a union-find over region ids.
//
// local: (2 rows + 1) x (2 cols + 1) row-major boundary strengths; part: rows
// x cols row-major region ids in [0, totcc); ucm: like local.
void merge_regions(const double* local, const int32_t* part, int rows, int cols, int totcc, double* ucm) {
  const int txc = 2 * rows + 1, tyc = 2 * cols + 1;
  std::fill(ucm, ucm + size_t(txc) * tyc, 0.0);
  // Boundary elements, gathered in row-major order (sequential reads of the
  // region ids and the strengths), grouped by pair with a stable counting
  // sort, then sorted within each pair by coordinate (the reference's
  // column-major doubled-grid index), so that each pair's sums run over its
  // elements in coordinate order as the reference's sorted boundary lists
  // do.  Pairs are found through each region's short neighbour list.
  struct Pair {
    double energy, total, length;
    std::vector<std::pair<int32_t, int32_t>> runs;  // [begin, end) ranges of coord
    // the queue entries last pushed for this pair: their energy, the pair's
    // region ids then, and the smallest first region id among them
    double qe;
    int32_t qlo, qhi, minr1;
  };
  std::vector<Pair> pairs;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> adj(static_cast<size_t>(totcc));  // (neighbour, pair)
  std::vector<int32_t> el_pair, el_coord;
  std::vector<double> el_val;
  el_pair.reserve(size_t(rows) * cols / 2);
  el_coord.reserve(size_t(rows) * cols / 2);
  el_val.reserve(size_t(rows) * cols / 2);
  auto pair_of = [&](int32_t a, int32_t b) {
    for (const auto& e : adj[size_t(a)])
      if (e.first == b) return e.second;
    const int32_t id = int32_t(pairs.size());
    pairs.push_back(Pair{0.0, 0.0, 0.0, {}, 0.0, -1, -1, -1});
    adj[size_t(a)].emplace_back(b, id);
    adj[size_t(b)].emplace_back(a, id);
    return id;
  };
  for (int x = 0; x < rows; x++) {
    const int32_t* pr = part + size_t(x) * cols;
    const double* l1 = local + size_t(2 * x + 1) * tyc;  // doubled row 2x + 1
    const double* l2 = l1 + tyc;                         // doubled row 2x + 2
    for (int y = 0; y < cols; y++) {
      const int32_t a = pr[y];
      if (x + 1 < rows) {
        const int32_t b = pr[size_t(cols) + y];
        if (a != b) {
          el_pair.push_back(pair_of(std::min(a, b), std::max(a, b)));
          el_coord.push_back((2 * x + 2) + (2 * y + 1) * txc);
          el_val.push_back(l2[2 * y + 1]);
        }
      }
      if (y + 1 < cols) {
        const int32_t b = pr[y + 1];
        if (a != b) {
          el_pair.push_back(pair_of(std::min(a, b), std::max(a, b)));
          el_coord.push_back((2 * x + 1) + (2 * y + 2) * txc);
          el_val.push_back(l1[2 * y + 2]);
        }
      }
    }
  }
  std::vector<int32_t> coord(el_coord.size());
  {
    std::vector<int32_t> start(pairs.size() + 1, 0);
    for (int32_t id : el_pair) start[size_t(id) + 1]++;
    for (size_t k = 0; k < pairs.size(); k++) start[k + 1] += start[k];
    std::vector<std::pair<int32_t, double>> cv(el_coord.size());
    std::vector<int32_t> at(start.begin(), start.end() - 1);
    for (size_t e = 0; e < el_pair.size(); e++) cv[size_t(at[size_t(el_pair[e])]++)] = {el_coord[e], el_val[e]};
    std::vector<int32_t>().swap(el_pair);
    std::vector<int32_t>().swap(el_coord);
    std::vector<double>().swap(el_val);
#pragma omp parallel for schedule(dynamic, 256)
    for (long k = 0; k < long(pairs.size()); k++) {
      auto b0 = cv.begin() + start[size_t(k)], b1 = cv.begin() + start[size_t(k) + 1];
      std::sort(b0, b1, [](const std::pair<int32_t, double>& p, const std::pair<int32_t, double>& q) {
        return p.first < q.first;
      });
      double len = 0, tot = 0;
      for (auto it = b0; it != b1; ++it) {
        len++;
        tot += it->second;
        coord[size_t(it - cv.begin())] = it->first;
      }
      Pair& q = pairs[size_t(k)];
      q.length = len;
      q.total = tot;
      q.runs.emplace_back(start[size_t(k)], start[size_t(k) + 1]);
    }
  }
  std::vector<std::unordered_map<int32_t, int32_t>> nb(static_cast<size_t>(totcc));
  for (size_t r = 0; r < adj.size(); r++) {
    nb[r].reserve(adj[r].size());
    for (const auto& e : adj[r]) nb[r].emplace(e.first, e.second);
    std::vector<std::pair<int32_t, int32_t>>().swap(adj[r]);
  }
  struct Node {
    double energy;
    int r1, r2;
  };
  auto later = [](const Node& a, const Node& b) {
    return a.energy > b.energy || (a.energy == b.energy && a.r1 > b.r1) ||
           (a.energy == b.energy && a.r1 == b.r1 && a.r2 > b.r2);
  };
  std::vector<Node> heap_store;
  heap_store.reserve(coord.size() + size_t(totcc) * 4);
  std::priority_queue<Node, std::vector<Node>, decltype(later)> queue(later, std::move(heap_store));
  for (size_t k = 0; k < pairs.size(); k++) {
    Pair& q = pairs[k];
    q.energy = q.total / q.length;
  }
  // Pushing (d, r1, r2) for a pair that already has an entry with the same
  // energy and regions and a first id <= r1 changes nothing: that entry
  // pops first, and if it is stale so is this one.  Such pushes are skipped.
  auto push = [&](Pair& q, double d, int32_t r1, int32_t r2) {
    const int32_t lo = std::min(r1, r2), hi = std::max(r1, r2);
    if (q.qe == d && q.qlo == lo && q.qhi == hi) {
      if (q.minr1 <= r1) return;
      q.minr1 = r1;
    } else {
      q.qe = d;
      q.qlo = lo;
      q.qhi = hi;
      q.minr1 = r1;
    }
    queue.push(Node{d, r1, r2});
  };
  for (size_t r = 0; r < nb.size(); r++)
    for (const auto& e : nb[r])
      if (e.first > int32_t(r)) push(pairs[size_t(e.second)], pairs[size_t(e.second)].energy, int32_t(r), e.first);
  std::vector<int32_t> parent(static_cast<size_t>(totcc));
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int32_t r) {
    while (parent[size_t(r)] != r) {
      parent[size_t(r)] = parent[size_t(parent[size_t(r)])];
      r = parent[size_t(r)];
    }
    return r;
  };
  while (!queue.empty()) {
    const Node top = queue.top();
    queue.pop();
    if (find(top.r1) != top.r1 || find(top.r2) != top.r2) continue;
    auto f12 = nb[size_t(top.r1)].find(top.r2);
    if (f12 == nb[size_t(top.r1)].end() || pairs[size_t(f12->second)].energy != top.energy) continue;
    Pair& joined = pairs[size_t(f12->second)];
    const double saliency = joined.total / joined.length;
    for (const auto& r : joined.runs)
      for (int32_t k = r.first; k < r.second; k++) {
        const int32_t c = coord[size_t(k)];
        double& u = ucm[size_t(c % txc) * tyc + c / txc];
        if (u < saliency) u = saliency;
      }
    joined.runs.clear();
    const int32_t son = std::min(top.r1, top.r2), father = std::max(top.r1, top.r2);
    parent[size_t(son)] = father;
    auto& fmap = nb[size_t(father)];
    auto& smap = nb[size_t(son)];
    // every map loses an absorbed region's id when it is absorbed, so the
    // father's map holds active regions only once the son is gone
    fmap.erase(son);
    for (const auto& kv : smap) {
      const int32_t n = kv.first;
      if (n == father || find(n) != n) continue;
      Pair& q = pairs[size_t(kv.second)];
      auto& nmap = nb[size_t(n)];
      nmap.erase(son);
      auto fit = fmap.find(n);
      if (fit == fmap.end()) {
        // a new neighbour of the father: 0 + the son's sums
        q.total = 0.0 + q.total;
        q.length = 0.0 + q.length;
        fmap.emplace(n, kv.second);
        nmap[father] = kv.second;
      } else {
        Pair& f = pairs[size_t(fit->second)];
        f.total += q.total;
        f.length += q.length;
        if (f.runs.size() < q.runs.size()) f.runs.swap(q.runs);
        f.runs.insert(f.runs.end(), q.runs.begin(), q.runs.end());
        q.runs.clear();
      }
    }
    smap.clear();
    for (const auto& kv : fmap) {
      Pair& q = pairs[size_t(kv.second)];
      const double d = q.total / q.length;
      push(q, d, kv.first, father);
      q.energy = d;
    }
  }
  // complete the map on the even (pixel-corner) sites by the max of the
  // four neighbouring boundary elements (CompleteContourMap, :232-251)
  for (int x = 0; x < txc; x += 2)
    for (int y = 0; y < tyc; y += 2) {
      double m = 0.0;
      if (x + 1 < txc && m < ucm[size_t(x + 1) * tyc + y]) m = ucm[size_t(x + 1) * tyc + y];
      if (y + 1 < tyc && m < ucm[size_t(x) * tyc + y + 1]) m = ucm[size_t(x) * tyc + y + 1];
      if (x >= 1 && m < ucm[size_t(x - 1) * tyc + y]) m = ucm[size_t(x - 1) * tyc + y];
      if (y >= 1 && m < ucm[size_t(x) * tyc + y - 1]) m = ucm[size_t(x) * tyc + y - 1];
      ucm[size_t(x) * tyc + y] = m;
    }
}

}  // namespace

std::vector<int32_t> watershed8(const uint8_t* a, int rows, int cols) {
  const size_t N = size_t(rows) * cols;
  std::vector<size_t> start(257, 0);
  for (size_t i = 0; i < N; i++) start[size_t(a[i]) + 1]++;
  for (int l = 0; l < 256; l++) start[size_t(l) + 1] += start[size_t(l)];
  std::vector<uint32_t> order(N);
  for (size_t i = 0; i < N; i++) order[start[a[i]]++] = uint32_t(i);
  std::vector<int32_t> L(N, -1);
  int32_t colour = 1;
  for (size_t k = 0; k < N; k++) {
    const size_t i = order[k];
    const long x = long(i / size_t(cols)), y = long(i % size_t(cols));
    int32_t state = -1;
    for (long dx = -1; dx <= 1; dx++)
      for (long dy = -1; dy <= 1; dy++) {
        if ((dx == 0 && dy == 0) || x + dx < 0 || x + dx >= rows || y + dy < 0 || y + dy >= cols) continue;
        const int32_t v = L[size_t(x + dx) * cols + (y + dy)];
        if (v > 0) state = (state == -1) ? v : (v != state ? 0 : state);
      }
    L[i] = state >= 0 ? state : colour++;
  }
  return L;
}

Contours fit_contours(const uint8_t* bw, int rows, int cols) {
  Builder b;
  b.sx = rows;
  b.sy = cols;
  b.bw = bw;
  b.build();
  b.subdivide_local();
  if (!b.subdivide_global()) {
    Contours empty;
    empty.failed = true;
    return empty;
  }
  return std::move(b.c);
}

void pb8_from_maps(const double* pbo, int rows, int cols, UcmStages& S) {
  // CreateFinestPartition, src/contours2ucm.cpp:255-278
  const auto t0 = std::chrono::steady_clock::now();
  const size_t N = size_t(rows) * cols;
  std::vector<double> pb(pbo, pbo + N);
  for (int k = 1; k < 8; k++)
    for (size_t i = 0; i < N; i++)
      if (pb[i] < pbo[size_t(k) * N + i]) pb[i] = pbo[size_t(k) * N + i];
  double lo = pb[0], hi = pb[0];
  for (size_t i = 1; i < N; i++) {
    lo = std::min(lo, pb[i]);
    hi = std::max(hi, pb[i]);
  }
  // convertTo(CV_8U, alpha, beta): round half to even, saturate
  const double alpha = 255 / (hi - lo);
  const double beta = -255 * lo / (hi - lo);
  S.pb8.resize(N);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; i++) {
    const double r = std::nearbyint(pb[i] * alpha + beta);
    S.pb8[i] = !(r >= 0) ? uint8_t(0) : (r > 255 ? uint8_t(255) : uint8_t(r));
  }
  S.ms[0] = ms_since(t0);
}

void finest_partition(int rows, int cols, UcmStages& S, std::vector<uint8_t>& omask) {
  // CreateFinestPartition, src/contours2ucm.cpp:280-404
  const size_t N = size_t(rows) * cols;
  auto t0 = std::chrono::steady_clock::now();
  S.ws = watershed8(S.pb8.data(), rows, cols);
  S.ms[1] = ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  std::vector<uint8_t> bw(N);
  for (size_t i = 0; i < N; i++) bw[i] = S.ws[i] == 0;
  S.contours = fit_contours(bw.data(), rows, cols);
  const Contours& ct = S.contours;
  omask.assign(N, 0);
  for (size_t e = 0; e < ct.es.size(); e++) {
    const int v1 = ct.es[e], v2 = ct.ee[e];
    const int r1 = ct.vx[size_t(v1)], c1 = ct.vy[size_t(v1)], r2 = ct.vx[size_t(v2)], c2 = ct.vy[size_t(v2)];
    // the chord's angle, in the reference's float steps
    float ang = float(M_PI / 2);
    if (c1 != c2) ang = float(std::atan((r1 - r2) * 1.0 / (c1 - c2)));
    const float deg = float(double(ang * 180) / M_PI);
    int o;
    if (deg < -78.75 || deg >= 78.75)
      o = 0;
    else if (deg >= 56.25)
      o = 1;
    else if (deg >= 33.75)
      o = 2;
    else if (deg >= 11.25)
      o = 3;
    else if (deg >= -11.25)
      o = 4;
    else if (deg >= -33.75)
      o = 5;
    else if (deg >= -56.25)
      o = 6;
    else
      o = 7;
    const uint8_t bit = uint8_t(1u << o);
    for (size_t p = 0; p < ct.px[e].size(); p++) omask[size_t(ct.px[e][p]) * cols + ct.py[e][p]] |= bit;
    omask[size_t(r1) * cols + c1] |= bit;
    omask[size_t(r2) * cols + c2] |= bit;
  }
  S.ms[2] = ms_since(t0);
}

void gather_strength(const double* pbo, const uint8_t* omask, size_t n, std::vector<double>& ws_wt) {
  // each contour pixel: the max over its edges' orientation maps, from 0
  ws_wt.assign(n, 0.0);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) {
    const unsigned m = omask[i];
    if (!m) continue;
    double v = 0.0;
    for (int o = 0; o < 8; o++)
      if (m & (1u << o)) v = std::max(pbo[size_t(o) * n + i], v);
    ws_wt[i] = v;
  }
}

void contours2ucm(const double* pbo, int rows, int cols, bool double_size, double* ucm, UcmStages* stages) {
  UcmStages local_stages;
  UcmStages& S = stages ? *stages : local_stages;
  pb8_from_maps(pbo, rows, cols, S);
  std::vector<uint8_t> omask;
  finest_partition(rows, cols, S, omask);
  const auto t0 = std::chrono::steady_clock::now();
  gather_strength(pbo, omask.data(), size_t(rows) * cols, S.ws_wt);
  S.ms[3] = ms_since(t0);
  ucm_from_strength(rows, cols, double_size, ucm, S);
}

void ucm_from_strength(int rows, int cols, bool double_size, double* ucm, UcmStages& S) {
  const size_t N = size_t(rows) * cols;
  auto t0 = std::chrono::steady_clock::now();
  // ---- the doubled grid (SuperContour4C, :24-78): pixels at even sites,
  // the min of two neighbours between them, the last row and column copied
  const int R2 = 2 * rows, C2 = 2 * cols;
  std::vector<double> pb2(size_t(R2) * C2);
  const auto& w = S.ws_wt;
  // per pixel row: its even row (pixel, horizontal min) and odd row
  // (vertical min, zero); every element is written once
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; r++) {
    double* ev = pb2.data() + size_t(2 * r) * C2;
    double* od = ev + C2;
    const double* wr = w.data() + size_t(r) * cols;
    for (int c = 0; c < cols; c++) {
      ev[2 * c] = wr[c];
      ev[2 * c + 1] = (c != cols - 1) ? std::min(wr[c], wr[c + 1]) : 0.0;
      od[2 * c] = (r != rows - 1) ? std::min(wr[c], wr[c + size_t(cols)]) : 0.0;
      od[2 * c + 1] = 0.0;
    }
  }
  for (int j = 0; j < C2; j++) pb2[size_t(R2 - 1) * C2 + j] = pb2[size_t(R2 - 2) * C2 + j];
  for (int i = 0; i < R2; i++)
    pb2[size_t(i) * C2 + C2 - 1] = std::max(pb2[size_t(i) * C2 + C2 - 1], pb2[size_t(i) * C2 + C2 - 2]);
  // ---- cleaning (CleanWatersheds, :80-189): a zero with no zero among its 8
  // neighbours is joined to the weakest of its four diagonal directions.
  // Reads outside the grid (a zero on the second row or column) are taken at
  // the row-major flat position, as the reference's unchecked accesses land,
  // and as 0 before the first element.
  auto val = [&](long i, long j) -> double {
    const long f = i * C2 + j;
    return f < 0 ? 0.0 : pb2[size_t(f)];
  };
  auto ref = [&](long i, long j) -> double& { return pb2[size_t(i * C2 + j)]; };
  // isolated zeros, found per row in parallel, then in raster order
  std::vector<std::vector<long>> art_rows(static_cast<size_t>(R2));
#pragma omp parallel for schedule(static)
  for (long i = 0; i < R2; i++)
    for (long j = 0; j < C2; j++) {
      if (pb2[size_t(i) * C2 + j] != 0) continue;
      bool isolated = true;
      for (long di = -1; di <= 1 && isolated; di++)
        for (long dj = -1; dj <= 1; dj++) {
          if ((di == 0 && dj == 0) || i + di < 0 || i + di >= R2 || j + dj < 0 || j + dj >= C2) continue;
          if (pb2[size_t(i + di) * C2 + j + dj] == 0) {
            isolated = false;
            break;
          }
        }
      if (isolated) art_rows[size_t(i)].push_back(i * C2 + j);
    }
  std::vector<long> art;
  for (const auto& r : art_rows) art.insert(art.end(), r.begin(), r.end());
  for (long a : art) {
    const long xc = a / C2, yc = a % C2;
    double vec[4];
    vec[0] = std::max(val(xc - 2, yc - 1), val(xc - 1, yc - 2));
    vec[1] = std::max(val(xc + 2, yc - 1), val(xc + 1, yc - 2));
    vec[2] = std::max(val(xc + 2, yc + 1), val(xc + 1, yc + 2));
    vec[3] = std::max(val(xc - 2, yc + 1), val(xc - 1, yc + 2));
    double nd = vec[0];
    int id = 0;
    for (int k = 1; k < 4; k++)
      if (vec[k] < nd) {
        id = k;
        nd = vec[k];
      }
    const long sxd = (id == 1 || id == 2) ? 1 : -1;  // row direction
    const long syd = (id >= 2) ? 1 : -1;             // column direction
    if (val(xc + 2 * sxd, yc + syd) < val(xc + sxd, yc + 2 * syd)) {
      ref(xc, yc + syd) = 0;
      ref(xc + sxd, yc) = vec[id];
    } else {
      ref(xc, yc + syd) = vec[id];
      ref(xc + sxd, yc) = 0;
    }
    ref(xc + sxd, yc + syd) = vec[id];
  }
  // ---- regions: 8-connected zeros of the cleaned grid, sampled at the odd
  // sites (one per pixel), numbered from 0
  S.labels.resize(N);
  const int ncomp = region_labels(pb2.data(), R2, C2, rows, cols, S.labels.data());
  int totcc = -1;
  for (size_t i = 0; i < N; i++) totcc = std::max(totcc, int(S.labels[i]));
  (void)ncomp;
  if (*std::min_element(S.labels.begin(), S.labels.end()) < 0)
    throw std::runtime_error("a pixel site of the cleaned grid is a boundary (the reference's region ids go negative)");
  totcc++;
  S.regions = totcc;
  // (2 rows + 1) x (2 cols + 1) with the last row and column repeated
  const int RR = R2 + 1, CC = C2 + 1;
  S.ws_wt2.resize(size_t(RR) * CC);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < RR; i++)
    for (int j = 0; j < CC; j++)
      S.ws_wt2[size_t(i) * CC + j] = pb2[size_t(i == R2 ? R2 - 1 : i) * C2 + (j == C2 ? C2 - 1 : j)];
  S.ms[4] = ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  // ---- merging (UcmMeanPb, src/ucm_mean_pb.cpp:402-449)
  S.super_ucm.resize(size_t(RR) * CC);
  merge_regions(S.ws_wt2.data(), S.labels.data(), rows, cols, totcc, S.super_ucm.data());
  S.ms[5] = ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  // ---- calibration (NormalizeOutput, :191-249) and the output size
  auto calib = [](double v) {
    const double comp = 1 * -2.7487 + v * 11.1189;
    double p = 1.0 / (1.0 + std::exp(-comp));
    p = (p - 0.0602) / (1.0 - 0.0602);
    return std::min(1.0, std::max(0.0, p));
  };
  if (double_size) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(RR) * CC; i++) ucm[i] = calib(S.super_ucm[size_t(i)]);
  } else {
    const int ro = (RR + 1) / 2, co = (CC + 1) / 2;
    std::fill(ucm, ucm + size_t(ro) * co, 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 2; i < RR; i += 2)
      for (int j = 2; j < CC; j += 2) ucm[size_t((i - 2) / 2) * co + (j - 2) / 2] = calib(S.super_ucm[size_t(i) * CC + j]);
  }
  S.ms[6] = ms_since(t0);
}

}  // namespace owt
}  // namespace gpb

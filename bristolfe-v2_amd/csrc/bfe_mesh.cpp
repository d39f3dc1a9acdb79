// Host mesh preprocessing (C++, multi-threaded) for multi-million element
// models: the geometric primitives under the reference's model builders.
//
//   bfe_element_centres          code/fn_calc_element_centres.m: mean of each
//                                element's nodes (absent nodes skipped)
//   bfe_points_in_polygon        MATLAB inpolygon as code/fn_elements_in_region.m
//                                uses it: inside or on the boundary
//   bfe_dist_points_to_polygon   code/fn_dist_point_to_bdry_2D.m: distance
//                                to the nearest vertex and to the two edges
//                                at that vertex (when the foot of the
//                                perpendicular lies on the edge)
//   bfe_tri_grid                 the node grid and triangles of
//                                code/fn_rectangular_structured_mesh.m:22-43
//                                and code/fn_isometric_structured_mesh.m:22-75
//                                before the region clip
//   bfe_remove_unused_nodes      code/internal/fn_remove_unused_nodes.m:
//                                renumbering of the nodes that elements use
//   bfe_get_edges                code/fn_get_edges.m:1-35: unique element
//                                edges, sorted, with the elements on each
//                                (what fn_find_free_edges.m and
//                                fn_add_fluid_solid_interface_els.m build on)
//
// Each function evaluates the same expressions in the same order as the
// numpy restatements in mesh.py (no fused multiply-add on the host), so the
// results are identical to the last bit.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <exception>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bfe.h"

namespace {

template <typename F>
void par_range(int64_t count, F&& fn) {
  unsigned hc = std::thread::hardware_concurrency();
  const int nw = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(hc ? hc : 4, 16)));
  constexpr int64_t kChunk = 1 << 14;
  if (nw == 1 || count < 2 * kChunk) {
    fn(0, count);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  for (int w = 0; w < nw; ++w)
    pool.emplace_back([&] {
      for (;;) {
        const int64_t i0 = next.fetch_add(kChunk);
        if (i0 >= count) break;
        fn(i0, std::min(i0 + kChunk, count));
      }
    });
  for (auto& t : pool) t.join();
}

// fn(i) for i in [0, count) on the worker threads, one item at a time (for
// a few large items, such as the buckets of a sort).
template <typename F>
void par_items(int64_t count, F&& fn) {
  unsigned hc = std::thread::hardware_concurrency();
  const int nw = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(hc ? hc : 4, 16)));
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  for (int w = 0; w < nw; ++w)
    pool.emplace_back([&] {
      for (int64_t i; (i = next.fetch_add(1)) < count;) fn(i);
    });
  for (auto& t : pool) t.join();
}

}  // namespace

extern "C" {

bfe_status bfe_element_centres(int64_t n_nodes, int32_t dim, const double* nds, int64_t n_el,
                               int32_t nper, const int64_t* els, double* out) {
  if (n_nodes < 0 || n_el < 0 || dim < 1 || dim > 3 || nper < 1 || (n_el && (!els || !out)) ||
      (n_nodes && !nds))
    return BFE_ERR_ARG;
  std::atomic<bool> bad{false};
  par_range(n_el * nper, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i)
      if (els[i] < 0 || els[i] > n_nodes) bad = true;
  });
  if (bad) return BFE_ERR_ARG;
  par_range(n_el, [&](int64_t e0, int64_t e1) {
    for (int64_t e = e0; e < e1; ++e) {
      const int64_t* v = els + e * nper;  // 1-based, 0 = absent
      int cnt = 0;
      for (int a = 0; a < nper; ++a) cnt += v[a] > 0;
      for (int c = 0; c < dim; ++c) {
        double s = 0.0;
        for (int a = 0; a < nper; ++a)
          if (v[a] > 0) s += nds[(v[a] - 1) * dim + c];
        out[e * dim + c] = cnt > 0 ? s / cnt : std::numeric_limits<double>::quiet_NaN();
      }
    }
  });
  return BFE_OK;
}

bfe_status bfe_points_in_polygon(int64_t n_pts, const double* pts, int32_t n_v,
                                 const double* vx, const double* vy, uint8_t* out) {
  if (n_pts < 0 || n_v < 1 || (n_pts && (!pts || !out)) || !vx || !vy) return BFE_ERR_ARG;
  double xmin = vx[0], xmax = vx[0], ymin = vy[0], ymax = vy[0];
  for (int32_t k = 1; k < n_v; ++k) {
    xmin = std::min(xmin, vx[k]);
    xmax = std::max(xmax, vx[k]);
    ymin = std::min(ymin, vy[k]);
    ymax = std::max(ymax, vy[k]);
  }
  const double span = std::max(std::max(xmax - xmin, ymax - ymin), 1e-300);
  const double eps = 1e-12 * span;
  par_range(n_pts, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
      const double x = pts[2 * i], y = pts[2 * i + 1];
      if (!(x >= xmin - eps && x <= xmax + eps && y >= ymin - eps && y <= ymax + eps)) {
        out[i] = 0;
        continue;
      }
      bool inside = false, on = false;
      for (int32_t k = 0; k < n_v; ++k) {
        const double x1 = vx[k], y1 = vy[k];
        const double x2 = vx[(k + 1) % n_v], y2 = vy[(k + 1) % n_v];
        const bool cond = (y1 > y) != (y2 > y);
        if (y2 != y1) {
          const double xint = x1 + (y - y1) * ((x2 - x1) / (y2 - y1));
          inside ^= cond && (x < xint);
        }
        const double cross = (x2 - x1) * (y - y1) - (y2 - y1) * (x - x1);
        const double scale = std::max(std::max(std::fabs(x2 - x1), std::fabs(y2 - y1)), 1e-300);
        const bool within =
            x >= std::min(x1, x2) - 1e-12 * scale && x <= std::max(x1, x2) + 1e-12 * scale &&
            y >= std::min(y1, y2) - 1e-12 * scale && y <= std::max(y1, y2) + 1e-12 * scale;
        on = on || (within && std::fabs(cross) <= 1e-12 * scale * scale);
      }
      out[i] = (inside || on) ? 1 : 0;
    }
  });
  return BFE_OK;
}

bfe_status bfe_dist_points_to_polygon(int64_t n_pts, const double* pts, int32_t n_b,
                                      const double* bdry, double* out) {
  if (n_pts < 0 || n_b < 2 || (n_pts && (!pts || !out)) || !bdry) return BFE_ERR_ARG;
  par_range(n_pts, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
      const double px = pts[2 * i], py = pts[2 * i + 1];
      int32_t j = 0;
      double r0 = std::numeric_limits<double>::infinity();
      for (int32_t k = 0; k < n_b; ++k) {
        const double dx = px - bdry[2 * k], dy = py - bdry[2 * k + 1];
        const double r = std::sqrt(dx * dx + dy * dy);
        if (r < r0) {  // first of equal minima, as argmin
          r0 = r;
          j = k;
        }
      }
      auto line = [&](int32_t a, int32_t b) {
        const double p1x = bdry[2 * a], p1y = bdry[2 * a + 1];
        double nx = bdry[2 * b] - p1x, ny = bdry[2 * b + 1] - p1y;
        const double d = std::sqrt(nx * nx + ny * ny);
        nx = nx / d;
        ny = ny / d;
        const double qx = px - p1x, qy = py - p1y;
        double s = qx * nx + qy * ny;
        const double rx = qx - s * nx, ry = qy - s * ny;
        s = s / d;
        const double r = std::sqrt(rx * rx + ry * ry);
        return (s < 0.0 || s > 1.0) ? std::numeric_limits<double>::infinity() : r;
      };
      const double r1 = line(j, (j + 1) % n_b);
      const double r2 = line(j, (j - 1 + n_b) % n_b);
      out[i] = std::min(std::min(r0, r1), r2);
    }
  });
  return BFE_OK;
}

// Node grid (node k = iy + ny ix at x[ix], y[iy]; on the isometric grid
// even rows (iy = 0, 2, ...) moved by +shift and odd rows by -shift, with
// shift = el_size / 4) and its triangles before the region clip, in the
// reference's order: blocks of element node triples, each block raveled
// column-major over (row, column). Node numbers are 1-based. With nds and
// els NULL only *n_el is set. Rectangular (iso = 0), blocks
//   a (iy, ix), (iy+1, ix+1), (iy+1, ix)   b (iy, ix), (iy, ix+1), (iy+1, ix+1)
// over iy < ny-1; isometric (iso = 1), a and b over even iy < ny-1 and
//   c (iy, ix+1), (iy+1, ix+1), (iy+1, ix)   d (iy, ix), (iy, ix+1), (iy+1, ix)
// over odd iy < ny-1.
bfe_status bfe_tri_grid(int32_t iso, int64_t nx, int64_t ny, const double* x, const double* y,
                        double shift, double* nds, int64_t* els, int64_t* n_el) {
  if (!n_el || nx < 2 || ny < 2 || (iso != 0 && iso != 1) || nx > (int64_t(1) << 31) ||
      ny > (int64_t(1) << 31))
    return BFE_ERR_ARG;
  const int64_t ncol = nx - 1;
  // rows of each block: all rows below the last (rectangular); even and odd
  // rows below the last (isometric)
  const int64_t r_all = ny - 1, r_even = ny / 2, r_odd = (ny - 1) / 2;
  std::array<int64_t, 4> rows{};
  if (iso) rows = {r_even, r_even, r_odd, r_odd};
  else rows = {r_all, r_all, 0, 0};
  int64_t total = 0;
  for (int64_t r : rows) total += r * ncol;
  *n_el = total;
  if (nds == nullptr && els == nullptr) return BFE_OK;
  if (!nds || !els || !x || !y) return BFE_ERR_ARG;
  par_range(nx * ny, [&](int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; ++k) {
      const int64_t ix = k / ny, iy = k - ix * ny;
      double xv = x[ix];
      if (iso) xv = (iy % 2 == 0) ? xv + shift : xv - shift;
      nds[2 * k] = xv;
      nds[2 * k + 1] = y[iy];
    }
  });
  auto nn = [&](int64_t iy, int64_t ix) { return iy + ny * ix + 1; };
  int64_t base = 0;
  for (int blk = 0; blk < 4; ++blk) {
    const int64_t nr = rows[blk];
    if (nr == 0) continue;
    par_range(nr * ncol, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) {
        const int64_t ix = i / nr, ir = i - ix * nr;
        // first row of the triangle's lower pair of nodes
        const int64_t iy = iso ? (blk < 2 ? 2 * ir : 2 * ir + 1) : ir;
        int64_t* e = els + 3 * (base + i);
        switch (blk) {
          case 0: e[0] = nn(iy, ix); e[1] = nn(iy + 1, ix + 1); e[2] = nn(iy + 1, ix); break;
          case 1: e[0] = nn(iy, ix); e[1] = nn(iy, ix + 1); e[2] = nn(iy + 1, ix + 1); break;
          case 2: e[0] = nn(iy, ix + 1); e[1] = nn(iy + 1, ix + 1); e[2] = nn(iy + 1, ix); break;
          default: e[0] = nn(iy, ix); e[1] = nn(iy, ix + 1); e[2] = nn(iy + 1, ix); break;
        }
      }
    });
    base += nr * ncol;
  }
  return BFE_OK;
}

// Renumbers the nodes used by the elements (1-based, 0 absent) in place:
// new numbers in increasing order of the old ones. old_nds (size n_nodes)
// receives the old number of each new node, *n_used their count.
bfe_status bfe_remove_unused_nodes(int64_t n_nodes, int64_t n_el, int32_t nper, int64_t* els,
                                   int64_t* old_nds, int64_t* n_used) {
  if (n_nodes < 0 || n_el < 0 || nper < 1 || !n_used || (n_el && !els) || (n_nodes && !old_nds))
    return BFE_ERR_ARG;
  const int64_t m = n_el * nper;
  std::atomic<bool> bad{false};
  std::vector<uint8_t> used(static_cast<size_t>(n_nodes) + 1, 0);
  par_range(m, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
      const int64_t v = els[i];
      if (v < 0 || v > n_nodes) bad = true;
      else if (v > 0) used[v] = 1;  // idempotent stores of 1
    }
  });
  if (bad) return BFE_ERR_ARG;
  std::vector<int64_t> nw(static_cast<size_t>(n_nodes) + 1, 0);
  int64_t c = 0;
  for (int64_t v = 1; v <= n_nodes; ++v)
    if (used[v]) {
      nw[v] = ++c;
      old_nds[c - 1] = v;
    }
  *n_used = c;
  par_range(m, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) els[i] = nw[els[i]];
  });
  return BFE_OK;
}

// Unique edges of the elements (consecutive non-zero nodes of each row, the
// last closing to the first), each as (lower, higher) node, in increasing
// lexicographic order, and the elements on each edge (1-based element
// numbers, largest first) as a list: edge i's elements are
// el[ptr[i] .. ptr[i+1]). Outputs hold n_el * nper entries at most (ed: 2
// per edge, ptr: one more). An element of k nodes gives k edges, so a
// 2-node element lists its edge twice.
bfe_status bfe_get_edges(int64_t n_el, int32_t nper, const int64_t* els, int64_t* n_ed,
                         int64_t* ed, int64_t* ptr, int64_t* el) {
  if (n_el < 0 || nper < 1 || nper > 64 || !n_ed || (n_el && (!els || !ed || !ptr || !el)))
    return BFE_ERR_ARG;
  struct Occ {
    int64_t lo, hi, e;
  };
  // occurrences, element by element (count, scan, fill)
  std::vector<int64_t> cnt(static_cast<size_t>(n_el) + 1, 0);
  std::atomic<bool> bad{false};
  int64_t vmax = 0;
  for (int64_t e = 0; e < n_el; ++e) {
    int k = 0;
    for (int a = 0; a < nper; ++a) {
      const int64_t v = els[e * nper + a];
      if (v < 0) bad = true;
      if (v > 0) { ++k; vmax = std::max(vmax, v); }
    }
    cnt[e + 1] = k >= 2 ? k : 0;  // a single node has no edge
  }
  if (bad) return BFE_ERR_ARG;
  for (int64_t e = 0; e < n_el; ++e) cnt[e + 1] += cnt[e];
  const int64_t m = cnt[n_el];
  std::vector<Occ> occ(static_cast<size_t>(m));
  par_range(n_el, [&](int64_t e0, int64_t e1) {
    int64_t v[64];
    for (int64_t e = e0; e < e1; ++e) {
      int k = 0;
      for (int a = 0; a < nper; ++a)
        if (els[e * nper + a] > 0) v[k++] = els[e * nper + a];
      if (k < 2) continue;
      int64_t o = cnt[e];
      for (int a = 0; a < k; ++a) {
        const int64_t p = v[a], q = v[(a + 1) % k];
        occ[o++] = {std::min(p, q), std::max(p, q), e + 1};
      }
    }
  });
  // bucket by the lower node (ranges of node numbers), then sort each bucket
  // by (lower, higher, element descending) on its own thread
  const int nb = 256;
  const int64_t span = vmax / nb + 1;
  std::vector<int64_t> bcnt(nb + 1, 0);
  for (const Occ& o : occ) ++bcnt[o.lo / span + 1];
  for (int b = 0; b < nb; ++b) bcnt[b + 1] += bcnt[b];
  std::vector<Occ> srt(static_cast<size_t>(m));
  {
    std::vector<int64_t> fill(bcnt.begin(), bcnt.end() - 1);
    for (const Occ& o : occ) srt[fill[o.lo / span]++] = o;
  }
  std::vector<Occ>().swap(occ);
  par_items(nb, [&](int64_t b) {
    std::sort(srt.begin() + bcnt[b], srt.begin() + bcnt[b + 1], [](const Occ& a, const Occ& c) {
      if (a.lo != c.lo) return a.lo < c.lo;
      if (a.hi != c.hi) return a.hi < c.hi;
      return a.e > c.e;
    });
  });
  int64_t ne = 0;
  for (int64_t i = 0; i < m; ++i) {
    if (i == 0 || srt[i].lo != srt[i - 1].lo || srt[i].hi != srt[i - 1].hi) {
      ed[2 * ne] = srt[i].lo;
      ed[2 * ne + 1] = srt[i].hi;
      ptr[ne] = i;
      ++ne;
    }
    el[i] = srt[i].e;
  }
  ptr[ne] = m;
  *n_ed = ne;
  return BFE_OK;
}

}  // extern "C"

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
//
// Each function evaluates the same expressions in the same order as the
// numpy restatements in mesh.py (no fused multiply-add on the host), so the
// results are identical to the last bit.
#include <algorithm>
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

}  // extern "C"

// Host-side FP64 set-up of the explicit march: coefficients, interface
// inverse and the SELL-64 operator layout. See bfe_plan.hpp.
#include "bfe_plan.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <utility>

namespace bfe {

namespace {

[[noreturn]] void bad(const std::string& msg) { throw std::invalid_argument(msg); }

int n_workers() {
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 4;
  return static_cast<int>(std::min<unsigned>(hc, 16));
}

template <typename F>
void parallel_for(int64_t count, F&& fn) {
  const int nw = static_cast<int>(std::min<int64_t>(n_workers(), std::max<int64_t>(count, 1)));
  if (nw <= 1 || count < 1024) {
    for (int64_t i = 0; i < count; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  std::exception_ptr err = nullptr;
  std::atomic<bool> failed{false};
  for (int w = 0; w < nw; ++w) {
    pool.emplace_back([&]() {
      try {
        for (;;) {
          int64_t i0 = next.fetch_add(256);
          if (i0 >= count || failed.load()) break;
          int64_t i1 = std::min<int64_t>(i0 + 256, count);
          for (int64_t i = i0; i < i1; ++i) fn(i);
        }
      } catch (...) {
        if (!failed.exchange(true)) err = std::current_exception();
      }
    });
  }
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace

Csr normalise_csr(int64_t n, const int64_t* ptr, const int32_t* col,
                  const double* val, const char* name) {
  Csr out;
  out.n = n;
  out.ptr.assign(n + 1, 0);
  if (ptr == nullptr) return out;  // empty matrix
  if (ptr[0] != 0) bad(std::string(name) + ": row pointer must start at 0");
  for (int64_t i = 0; i < n; ++i)
    if (ptr[i + 1] < ptr[i]) bad(std::string(name) + ": row pointers must be non-decreasing");
  const int64_t nnz = ptr[n];
  if (nnz > 0 && (col == nullptr || val == nullptr))
    bad(std::string(name) + ": missing column or value array");
  std::vector<char> in_range(static_cast<size_t>(n), 1);
  parallel_for(n, [&](int64_t i) {
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p)
      if (col[p] < 0 || col[p] >= n) in_range[i] = 0;
  });
  for (int64_t i = 0; i < n; ++i)
    if (!in_range[i]) bad(std::string(name) + ": column index out of range");
  // First pass: count merged entries per row (sorted, duplicates summed,
  // exact zeros dropped, as MATLAB's sparse() stores them).
  std::vector<int64_t> new_len(n, 0);
  parallel_for(n, [&](int64_t i) {
    const int64_t b = ptr[i], e = ptr[i + 1];
    bool sorted = true;
    for (int64_t p = b + 1; p < e; ++p)
      if (col[p] <= col[p - 1]) { sorted = false; break; }
    int64_t len = 0;
    if (sorted) {
      for (int64_t p = b; p < e; ++p) len += (val[p] != 0.0);
    } else {
      std::vector<std::pair<int32_t, double>> r;
      r.reserve(e - b);
      for (int64_t p = b; p < e; ++p) r.emplace_back(col[p], val[p]);
      std::sort(r.begin(), r.end(), [](auto& a, auto& c) { return a.first < c.first; });
      for (size_t k = 0; k < r.size();) {
        size_t k2 = k;
        double s = 0.0;
        while (k2 < r.size() && r[k2].first == r[k].first) s += r[k2++].second;
        len += (s != 0.0);
        k = k2;
      }
    }
    new_len[i] = len;
  });
  for (int64_t i = 0; i < n; ++i) out.ptr[i + 1] = out.ptr[i] + new_len[i];
  out.col.resize(out.ptr[n]);
  out.val.resize(out.ptr[n]);
  parallel_for(n, [&](int64_t i) {
    const int64_t b = ptr[i], e = ptr[i + 1];
    int64_t o = out.ptr[i];
    std::vector<std::pair<int32_t, double>> r;
    r.reserve(e - b);
    for (int64_t p = b; p < e; ++p) r.emplace_back(col[p], val[p]);
    std::stable_sort(r.begin(), r.end(), [](auto& a, auto& c) { return a.first < c.first; });
    for (size_t k = 0; k < r.size();) {
      size_t k2 = k;
      double s = 0.0;
      while (k2 < r.size() && r[k2].first == r[k].first) s += r[k2++].second;
      if (s != 0.0) {
        out.col[o] = r[k].first;
        out.val[o] = s;
        ++o;
      }
      k = k2;
    }
  });
  for (double v : out.val)
    if (!std::isfinite(v)) bad(std::string(name) + ": non-finite entry");
  return out;
}

namespace {

// Dense LU with partial pivoting, solving A x = e_c in place (A is s x s,
// row-major). Returns false if a pivot is exactly zero.
bool dense_solve_unit(std::vector<double>& A, int s, int c, std::vector<double>& x) {
  std::vector<int> piv(s);
  for (int k = 0; k < s; ++k) {
    int pr = k;
    double best = std::fabs(A[(size_t)k * s + k]);
    for (int r = k + 1; r < s; ++r) {
      double v = std::fabs(A[(size_t)r * s + k]);
      if (v > best) { best = v; pr = r; }
    }
    if (best == 0.0) return false;
    piv[k] = pr;
    if (pr != k)
      for (int j = 0; j < s; ++j) std::swap(A[(size_t)k * s + j], A[(size_t)pr * s + j]);
    const double inv = 1.0 / A[(size_t)k * s + k];
    for (int r = k + 1; r < s; ++r) {
      double f = A[(size_t)r * s + k] * inv;
      if (f == 0.0) continue;
      A[(size_t)r * s + k] = f;
      for (int j = k + 1; j < s; ++j) A[(size_t)r * s + j] -= f * A[(size_t)k * s + j];
    }
  }
  x.assign(s, 0.0);
  x[c] = 1.0;
  for (int k = 0; k < s; ++k)
    if (piv[k] != k) std::swap(x[k], x[piv[k]]);
  for (int r = 1; r < s; ++r) {
    double acc = x[r];
    for (int j = 0; j < r; ++j) acc -= A[(size_t)r * s + j] * x[j];
    x[r] = acc;
  }
  for (int r = s - 1; r >= 0; --r) {
    double acc = x[r];
    for (int j = r + 1; j < s; ++j) acc -= A[(size_t)r * s + j] * x[j];
    x[r] = acc / A[(size_t)r * s + r];
  }
  return true;
}

}  // namespace

Csr truncated_inverse(const Csr& X, double drop_tol, const std::vector<double>& w) {
  const int64_t nI = X.n;
  if (!w.empty() && static_cast<int64_t>(w.size()) != nI)
    throw std::invalid_argument("truncated_inverse: weight count differs from the block size");
  auto wt = [&](int32_t g) { return w.empty() ? 1.0 : w[g]; };
  // symmetric adjacency of X's pattern (off-diagonal)
  std::vector<std::vector<int32_t>> adj(nI);
  for (int64_t i = 0; i < nI; ++i)
    for (int64_t p = X.ptr[i]; p < X.ptr[i + 1]; ++p) {
      int32_t j = X.col[p];
      if (j != i) {
        adj[i].push_back(j);
        adj[j].push_back(static_cast<int32_t>(i));
      }
    }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  constexpr int kMaxWindow = 4096;
  std::vector<std::vector<std::pair<int32_t, double>>> cols(nI);  // (row, value)
  std::atomic<bool> failed{false};
  std::string fail_msg;
  parallel_for(nI, [&](int64_t l) {
    if (failed.load()) return;
    std::vector<int32_t> order;   // BFS order (window members)
    std::vector<int32_t> depth;
    int radius = 6;
    for (;;) {
      // BFS to 'radius'; 'mark' holds the visited set (window is small)
      order.clear();
      depth.clear();
      std::vector<int32_t> mark;
      auto visited = [&](int32_t w) {
        auto it = std::lower_bound(mark.begin(), mark.end(), w);
        return it != mark.end() && *it == w;
      };
      auto visit = [&](int32_t w) { mark.insert(std::lower_bound(mark.begin(), mark.end(), w), w); };
      order.push_back(static_cast<int32_t>(l));
      depth.push_back(0);
      visit(static_cast<int32_t>(l));
      bool exhausted = true;
      for (size_t h = 0; h < order.size(); ++h) {
        int32_t v = order[h];
        int32_t d = depth[h];
        for (int32_t w : adj[v]) {
          if (visited(w)) continue;
          if (d + 1 > radius) { exhausted = false; continue; }
          visit(w);
          order.push_back(w);
          depth.push_back(d + 1);
        }
        if ((int)order.size() > kMaxWindow) break;
      }
      const int s = static_cast<int>(order.size());
      if (s > kMaxWindow) {
        if (!failed.exchange(true))
          fail_msg = "interface block too strongly coupled for a truncated inverse (window > 4096)";
        return;
      }
      // local index lookup
      std::vector<std::pair<int32_t, int32_t>> lut(s);
      for (int a = 0; a < s; ++a) lut[a] = {order[a], a};
      std::sort(lut.begin(), lut.end());
      auto loc = [&](int32_t g) -> int {
        auto it = std::lower_bound(lut.begin(), lut.end(), std::make_pair(g, -1));
        return (it != lut.end() && it->first == g) ? it->second : -1;
      };
      // the equilibrated block W X W^-1 (W = diag(w)): column l of X^-1 is
      // w_l W^-1 z with (W X W^-1) z = e_l, and z is what the tests weigh
      std::vector<double> A((size_t)s * s, 0.0);
      for (int a = 0; a < s; ++a) {
        int32_t g = order[a];
        for (int64_t p = X.ptr[g]; p < X.ptr[g + 1]; ++p) {
          int b = loc(X.col[p]);
          if (b >= 0) A[(size_t)a * s + b] += X.val[p] * (wt(g) / wt(X.col[p]));
        }
      }
      std::vector<double> x;
      if (!dense_solve_unit(A, s, 0, x)) {
        if (!failed.exchange(true)) fail_msg = "singular interface block X_II";
        return;
      }
      double xmax = 0.0, xbdry = 0.0;
      for (int a = 0; a < s; ++a) {
        const double xa = std::fabs(x[a]);
        xmax = std::max(xmax, xa);
        if (depth[a] == radius) xbdry = std::max(xbdry, xa);
      }
      if (!std::isfinite(xmax)) {
        if (!failed.exchange(true)) fail_msg = "non-finite interface inverse";
        return;
      }
      if (exhausted || xbdry <= 1e-20 * xmax) {
        auto& c = cols[l];
        for (int a = 0; a < s; ++a)
          if (std::fabs(x[a]) > drop_tol * xmax)
            c.emplace_back(order[a], x[a] * (wt(static_cast<int32_t>(l)) / wt(order[a])));
        return;
      }
      radius *= 2;
    }
  });
  if (failed.load()) throw std::domain_error(fail_msg);
  // transpose column lists into CSR rows
  Csr out;
  out.n = nI;
  out.ptr.assign(nI + 1, 0);
  for (int64_t l = 0; l < nI; ++l)
    for (auto& e : cols[l]) out.ptr[e.first + 1]++;
  for (int64_t i = 0; i < nI; ++i) out.ptr[i + 1] += out.ptr[i];
  out.col.resize(out.ptr[nI]);
  out.val.resize(out.ptr[nI]);
  std::vector<int64_t> fill(out.ptr.begin(), out.ptr.end() - 1);
  for (int64_t l = 0; l < nI; ++l)  // ascending l gives sorted columns per row
    for (auto& e : cols[l]) {
      int64_t o = fill[e.first]++;
      out.col[o] = static_cast<int32_t>(l);
      out.val[o] = e.second;
    }
  return out;
}

bool csr_is_normal(int64_t n, const int64_t* ptr, const int32_t* col, const double* val) {
  if (ptr == nullptr || ptr[0] != 0) return false;
  for (int64_t i = 0; i < n; ++i)
    if (ptr[i + 1] < ptr[i]) return false;
  if (ptr[n] > 0 && (col == nullptr || val == nullptr)) return false;
  std::vector<char> ok(static_cast<size_t>(n), 1);
  parallel_for(n, [&](int64_t i) {
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p)
      if (col[p] < 0 || col[p] >= n || (p > ptr[i] && col[p] <= col[p - 1]) ||
          !std::isfinite(val[p]) || val[p] == 0.0) {
        ok[i] = 0;
        return;
      }
  });
  for (int64_t i = 0; i < n; ++i)
    if (!ok[i]) return false;
  return true;
}

Plan build_plan(int64_t n, const CsrRef& K, const double* m, const Csr& C, double dt,
                int mode, bool host_sell) {
  if (n <= 0) bad("n must be positive");
  if (!(dt > 0.0) || !std::isfinite(dt)) bad("dt must be positive and finite");
  if (mode != 0 && mode != 1) bad("unknown solver mode");
  if (m == nullptr) bad("lumped mass is NULL");
  for (int64_t i = 0; i < n; ++i) {
    if (!std::isfinite(m[i])) bad("lumped mass has a non-finite entry");
    if (m[i] == 0.0)
      throw std::domain_error("lumped mass is zero at DOF " + std::to_string(i) +
                              " (M^-1 undefined)");
  }
  Plan P;
  P.n = n;
  P.mode = mode;
  P.dt = dt;
  P.nnz_k = K.nnz();

  // diagonal of C, interface set I (rows/cols touched by off-diagonal C)
  std::vector<double> cdiag(n, 0.0);
  std::vector<char> in_i(n, 0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = C.ptr[i]; p < C.ptr[i + 1]; ++p) {
      if (C.col[p] == i) cdiag[i] = C.val[p];
      else { in_i[i] = 1; in_i[C.col[p]] = 1; }
    }
  P.iface_pos.assign(n, -1);
  for (int64_t i = 0; i < n; ++i)
    if (in_i[i]) {
      P.iface_pos[i] = static_cast<int32_t>(P.iface.size());
      P.iface.push_back(static_cast<int32_t>(i));
    }

  P.alpha.resize(n);
  P.beta.resize(n);
  P.gamma.resize(n);
  bool damped = false;
  for (int64_t i = 0; i < n; ++i) {
    const double mi = m[i], ci = cdiag[i];
    double a, b, g;
    if (mode == 0 && !in_i[i]) {
      // X_ii = 1 + dt/2 c_i/m_i divides the whole row (v6:136-138)
      const double den = mi + 0.5 * dt * ci;
      if (den == 0.0) throw std::domain_error("singular diagonal of X");
      a = dt * dt / den;
      b = 2.0 * mi / den;
      g = -(mi - 0.5 * dt * ci) / den;
    } else if (mode == 0) {
      // interface rows: the right-hand side r_i; X_II is applied afterwards
      a = dt * dt / mi;
      b = 2.0;
      g = -1.0 + 0.5 * dt * ci / mi;
    } else {
      // fully explicit (v6:131-134)
      a = dt * dt / mi;
      b = 2.0 - dt * ci / mi;
      g = dt * ci / mi - 1.0;
    }
    P.alpha[i] = a;
    P.beta[i] = b;
    P.gamma[i] = g;
    if (b != 2.0) damped = true;
  }
  P.damped = damped;

  // SELL-64 of alpha-scaled K
  Sell64& S = P.op;
  S.n = n;
  S.n_slices = (n + 63) / 64;
  if (host_sell) {
  S.off.assign(S.n_slices + 1, 0);
  for (int64_t s = 0; s < S.n_slices; ++s) {
    int64_t w = 0;
    for (int64_t r = s * 64; r < std::min<int64_t>(n, s * 64 + 64); ++r)
      w = std::max<int64_t>(w, K.ptr[r + 1] - K.ptr[r]);
    S.off[s + 1] = S.off[s] + 64 * w;
  }
  S.col.resize(S.off[S.n_slices]);
  S.val.resize(S.off[S.n_slices]);
  parallel_for(S.n_slices, [&](int64_t s) {
    const int64_t w = (S.off[s + 1] - S.off[s]) / 64;
    for (int lane = 0; lane < 64; ++lane) {
      const int64_t r = s * 64 + lane;
      const int64_t rr = std::min<int64_t>(r, n - 1);
      const int64_t b = r < n ? K.ptr[r] : 0, e = r < n ? K.ptr[r + 1] : 0;
      for (int64_t k = 0; k < w; ++k) {
        const int64_t o = S.off[s] + 64 * k + lane;
        if (b + k < e) {
          S.col[o] = K.col[b + k];
          S.val[o] = P.alpha[r] * K.val[b + k];
        } else {
          S.col[o] = static_cast<int32_t>(rr);
          S.val[o] = 0.0;
        }
      }
    }
  });
  }

  // interface correction
  const int64_t nI = static_cast<int64_t>(P.iface.size());
  if (nI > 0) {
    P.coff.n = nI;
    P.coff.ptr.assign(nI + 1, 0);
    for (int64_t k = 0; k < nI; ++k) {
      const int64_t i = P.iface[k];
      for (int64_t p = C.ptr[i]; p < C.ptr[i + 1]; ++p)
        if (C.col[p] != i) {
          P.coff.col.push_back(C.col[p]);
          P.coff.val.push_back(C.val[p] / m[i]);
        }
      P.coff.ptr[k + 1] = static_cast<int64_t>(P.coff.col.size());
    }
    if (mode == 0) {
      P.sq = 0.0;
      P.sr = 0.5 * dt;
      // X_II = I + dt/2 M^-1 C restricted to I (diagonal c_i included)
      Csr X;
      X.n = nI;
      X.ptr.assign(nI + 1, 0);
      for (int64_t k = 0; k < nI; ++k) {
        const int64_t i = P.iface[k];
        bool diag_seen = false;
        for (int64_t p = C.ptr[i]; p < C.ptr[i + 1]; ++p) {
          const int32_t j = C.col[p];
          const int32_t l = P.iface_pos[j];
          double v = 0.5 * dt * C.val[p] / m[i];
          if (j == i) { v += 1.0; diag_seen = true; }
          X.col.push_back(l);
          X.val.push_back(v);
        }
        if (!diag_seen) {
          X.col.push_back(static_cast<int32_t>(k));
          X.val.push_back(1.0);
        }
        X.ptr[k + 1] = static_cast<int64_t>(X.col.size());
      }
      std::vector<double> w(nI);  // balance the block's very different row scales
      for (int64_t k = 0; k < nI; ++k) w[k] = std::sqrt(std::fabs(m[P.iface[k]]));
      P.xinv = truncated_inverse(X, 1e-18, w);
      P.has_xinv = true;
    } else {
      P.sq = -dt;
      P.sr = dt;
    }
    P.gfix.resize(nI);
    for (int64_t k = 0; k < nI; ++k) {
      const int64_t i = P.iface[k];
      P.gfix[k] = P.gamma[i] - (1.0 - P.beta[i]);
    }
  }
  return P;
}

}  // namespace bfe

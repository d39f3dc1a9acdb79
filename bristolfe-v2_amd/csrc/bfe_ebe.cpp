// Host-side layout of the element-by-element stiffness operator.
#include "bfe_ebe.hpp"

#include <algorithm>
#include <atomic>
#include <numeric>
#include <string>
#include <stdexcept>
#include <thread>

namespace bfe {
namespace {

template <typename F>
void par_nodes(int64_t count, F&& fn) {
  unsigned hc = std::thread::hardware_concurrency();
  const int nw = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(hc ? hc : 4, 16)));
  if (nw == 1 || count < 65536) {
    fn(0, count);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t chunk = (count + nw - 1) / nw;
  for (int w = 0; w < nw; ++w) {
    const int64_t b = w * chunk, e = std::min<int64_t>(count, b + chunk);
    if (b < e) pool.emplace_back([&fn, b, e] { fn(b, e); });
  }
  for (auto& t : pool) t.join();
}

}  // namespace

EbeLayout build_ebe_layout(int64_t n_nodes, int32_t dpn, int64_t n_el, int32_t npe,
                           const int32_t* el_nodes, const int32_t* el_tmpl, int32_t n_tmpl,
                           const double* tmpl, int32_t pad, int32_t max_slots) {
  EbeLayout L;
  L.npe = npe;
  L.dpn = dpn;
  L.n_nodes = n_nodes;
  const int KE = npe * dpn;
  // slots: the (template, vertex) pairs that occur, in (template, vertex) order
  std::vector<int32_t> slot_of(static_cast<size_t>(n_tmpl) * npe, -1);
  for (int64_t e = 0; e < n_el; ++e)
    for (int l = 0; l < npe; ++l) slot_of[static_cast<size_t>(el_tmpl[e]) * npe + l] = 0;
  std::vector<int32_t> pair;
  for (size_t i = 0; i < slot_of.size(); ++i)
    if (slot_of[i] == 0) {
      slot_of[i] = static_cast<int32_t>(pair.size());
      pair.push_back(static_cast<int32_t>(i));
    }
  const int32_t ns = static_cast<int32_t>(pair.size());
  if (ns > max_slots)
    throw std::invalid_argument("element operator: more than " + std::to_string(max_slots) +
                                " (template, vertex) slots; use the CSR operator");
  std::vector<int32_t> adj0(static_cast<size_t>(ns) * n_nodes, -1);
  for (int64_t e = 0; e < n_el; ++e)
    for (int l = 0; l < npe; ++l) {
      const int32_t v = el_nodes[e * npe + l];
      int32_t& a = adj0[static_cast<size_t>(slot_of[static_cast<size_t>(el_tmpl[e]) * npe + l]) *
                            n_nodes + v];
      if (a != -1)
        throw std::invalid_argument("element operator: two elements of one template share a "
                                    "node at the same local vertex; use the CSR operator");
      a = static_cast<int32_t>(e);
    }

  // neighbour classes, from an interior node (every slot occupied)
  int64_t ref = -1;
  for (int64_t v = 0; v < n_nodes && ref < 0; ++v) {
    int32_t c = 0;
    for (int32_t k = 0; k < ns; ++k) c += adj0[static_cast<size_t>(k) * n_nodes + v] >= 0;
    if (c == ns) ref = v;
  }
  std::vector<int32_t> cls0;
  int32_t ncls = 0;
  std::vector<int32_t> nbr;
  if (ref >= 0) {
    std::vector<int64_t> off(static_cast<size_t>(ns) * npe);
    for (int32_t k = 0; k < ns; ++k) {
      const int32_t e = adj0[static_cast<size_t>(k) * n_nodes + ref];
      for (int b = 0; b < npe; ++b) off[k * npe + b] = static_cast<int64_t>(el_nodes[e * npe + b]) - ref;
    }
    std::vector<int64_t> dist(off);
    std::sort(dist.begin(), dist.end());
    dist.erase(std::unique(dist.begin(), dist.end()), dist.end());
    ncls = static_cast<int32_t>(dist.size());
    cls0.resize(off.size());
    for (size_t i = 0; i < off.size(); ++i)
      cls0[i] = static_cast<int32_t>(std::lower_bound(dist.begin(), dist.end(), off[i]) - dist.begin());
    L.self_cls = static_cast<int32_t>(std::lower_bound(dist.begin(), dist.end(), 0) - dist.begin());
    L.cls_off = dist;
    // every node: the vertices of one class must be one node
    nbr.assign(static_cast<size_t>(ncls) * n_nodes, -1);
    std::atomic<bool> bad{false};
    par_nodes(n_nodes, [&](int64_t v0, int64_t v1) {
      for (int64_t v = v0; v < v1 && !bad.load(std::memory_order_relaxed); ++v)
        for (int32_t k = 0; k < ns; ++k) {
          const int32_t e = adj0[static_cast<size_t>(k) * n_nodes + v];
          if (e < 0) continue;
          for (int b = 0; b < npe; ++b) {
            int32_t& slotv = nbr[static_cast<size_t>(cls0[k * npe + b]) * n_nodes + v];
            const int32_t w = el_nodes[static_cast<int64_t>(e) * npe + b];
            if (slotv == -1) slotv = w;
            else if (slotv != w) bad.store(true, std::memory_order_relaxed);
          }
        }
    });
    if (bad.load()) {
      ncls = 0;
      cls0.clear();
      nbr.clear();
      L.self_cls = -1;
      L.cls_off.clear();
    }
  }

  // canonical orders: vertices of a slot by class, slots by their class list
  std::vector<std::vector<int32_t>> perm(ns), key(ns);
  for (int32_t k = 0; k < ns; ++k) {
    perm[k].resize(npe);
    std::iota(perm[k].begin(), perm[k].end(), 0);
    if (ncls > 0) {
      std::stable_sort(perm[k].begin(), perm[k].end(),
                       [&](int a, int b) { return cls0[k * npe + a] < cls0[k * npe + b]; });
      for (int b = 0; b < npe; ++b) key[k].push_back(cls0[k * npe + perm[k][b]]);
    }
  }
  std::vector<int32_t> order(ns);
  std::iota(order.begin(), order.end(), 0);
  if (ncls > 0)
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });

  L.n_slot = ns;
  L.n_slot_padded = (ns + pad - 1) / pad * pad;
  L.adj.assign(static_cast<size_t>(L.n_slot_padded) * n_nodes, -1);
  L.srows.assign(static_cast<size_t>(L.n_slot_padded) * dpn * KE, 0.0);
  for (int32_t kn = 0; kn < ns; ++kn) {
    const int32_t ko = order[kn];
    std::copy(adj0.begin() + static_cast<size_t>(ko) * n_nodes,
              adj0.begin() + static_cast<size_t>(ko + 1) * n_nodes,
              L.adj.begin() + static_cast<size_t>(kn) * n_nodes);
    const int32_t tau = pair[ko] / npe, l = pair[ko] % npe;
    for (int c = 0; c < dpn; ++c)
      for (int j = 0; j < KE; ++j)
        L.srows[(static_cast<size_t>(kn) * dpn + c) * KE + j] =
            tmpl[(static_cast<size_t>(tau) * KE + l * dpn + c) * KE + j];
  }
  if (ncls > 0 && ns <= 31) {  // the per-node slot mask has 31 slot bits
    L.n_cls = ncls;
    L.present.assign(n_nodes, 0u);
    par_nodes(n_nodes, [&](int64_t v0, int64_t v1) {
      for (int64_t v = v0; v < v1; ++v) {
        uint32_t m = 0;
        for (int32_t kn = 0; kn < ns; ++kn)
          if (L.adj[static_cast<size_t>(kn) * n_nodes + v] >= 0) m |= 1u << kn;
        for (int32_t j = 0; j < ncls; ++j) {
          const int32_t w = nbr[static_cast<size_t>(j) * n_nodes + v];
          if (w >= 0 && w != v + L.cls_off[j]) m |= 1u << 31;
        }
        L.present[v] = m;
      }
    });
    L.nbr = std::move(nbr);
    L.cls.resize(static_cast<size_t>(ns) * npe);
    L.srows_c.assign(L.srows.size(), 0.0);
    for (int32_t kn = 0; kn < ns; ++kn) {
      const int32_t ko = order[kn];
      for (int bn = 0; bn < npe; ++bn) {
        const int bo = perm[ko][bn];
        L.cls[kn * npe + bn] = cls0[ko * npe + bo];
        for (int c = 0; c < dpn; ++c)
          for (int d = 0; d < dpn; ++d)
            L.srows_c[(static_cast<size_t>(kn) * dpn + c) * KE + bn * dpn + d] =
                L.srows[(static_cast<size_t>(kn) * dpn + c) * KE + bo * dpn + d];
      }
    }
  }
  L.nodes4.resize(static_cast<size_t>(n_el) * 4);
  for (int64_t e = 0; e < n_el; ++e)
    for (int l = 0; l < 4; ++l) L.nodes4[e * 4 + l] = el_nodes[e * npe + (l < npe ? l : 0)];
  return L;
}

}  // namespace bfe

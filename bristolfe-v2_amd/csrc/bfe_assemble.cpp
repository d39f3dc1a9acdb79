// Native global assembly (host, FP64, multi-threaded) for large meshes.
//
// Same result as code/internal/fn_build_global_matrices_v5.m:50-159 for the
// element types CPE3, AC2D3 and C3D4 (no interface elements): element
// matrices with the absorbing modifications (:115-116), global index
// (node-1)*max_df + dof (:127-128), duplicates summed (:145-147), DOFs whose
// column is empty in K, M and C removed in order (fn_reduce_global_matrices.m)
// and gl_lookup (fn_create_fast_lookup.m). K is emitted row by row directly
// from a node-to-element adjacency: row r of node v sums the rows of K_e of
// the elements around v, so no coordinate list of n_el*k^2 entries is built.
// M and C are diagonal for these element types and are returned as vectors.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bfe.h"

namespace {

enum ElType : int32_t { kCPE3 = 0, kAC2D3 = 1, kC3D4 = 2 };

int nodes_of(int32_t t) { return t == kC3D4 ? 4 : 3; }

template <typename F>
void par_for(int64_t count, F&& fn) {
  unsigned hc = std::thread::hardware_concurrency();
  const int nw = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(hc ? hc : 4, 16)));
  if (nw == 1 || count < 4096) {
    for (int64_t i = 0; i < count; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::atomic<bool> failed{false};
  std::exception_ptr err = nullptr;
  std::vector<std::thread> pool;
  for (int w = 0; w < nw; ++w)
    pool.emplace_back([&] {
      try {
        for (;;) {
          const int64_t i0 = next.fetch_add(1024);
          if (i0 >= count || failed.load()) break;
          const int64_t i1 = std::min<int64_t>(i0 + 1024, count);
          for (int64_t i = i0; i < i1; ++i) fn(i);
        }
      } catch (...) {
        if (!failed.exchange(true)) err = std::current_exception();
      }
    });
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
}

struct Mesh {
  int64_t n_nodes = 0;
  int32_t dim = 0;
  const double* nds = nullptr;
  int64_t n_el = 0;
  int32_t nper = 0;
  const int32_t* els = nullptr;
  const int32_t* typ = nullptr;
  const int32_t* mat = nullptr;
  const double* abs_i = nullptr;
  int32_t n_mat = 0;
  const double* D = nullptr;    // n_mat x 36 (6x6 row-major); AC2D3 uses D[0] = bulk modulus
  const double* rho = nullptr;  // n_mat
  int32_t dof_mask = 0;         // bit d-1 set: DOF d in use (0 = all)
  double p_law = 3.0, max_damp = 3.1415e7, log_msr = std::log(0.01);
};

// Local DOFs of an element type after the dof_to_use filter (1-based DOF numbers).
void local_dofs(int32_t t, int32_t mask, std::vector<int>& ln, std::vector<int>& ld) {
  ln.clear();
  ld.clear();
  auto keep = [&](int d) { return mask == 0 || (mask >> (d - 1)) & 1; };
  if (t == kAC2D3) {
    for (int a = 0; a < 3; ++a)
      if (keep(4)) { ln.push_back(a); ld.push_back(4); }
  } else {
    const int nn = nodes_of(t);
    for (int a = 0; a < nn; ++a)
      for (int d = 1; d <= 3; ++d)
        if (keep(d)) { ln.push_back(a); ld.push_back(d); }
  }
}

// Full element stiffness (all DOFs of the type, node-major) and lumped mass
// per DOF. Closed forms as in elements.py: CPE3 Ke = (J/2) B'DB with
// engineering strains (e11,e22,e33,g23,g13,g12), d/dz = 0; AC2D3
// Ke = -(b b' + c c')/(2 J rho), M = -J/(6K); C3D4 Ke = (J/6) B'DB, M = J rho/24.
int element_matrix(const Mesh& m, int64_t e, double* K, double* mdiag) {
  const int32_t t = m.typ[e];
  const int32_t* v = m.els + e * m.nper;
  const double* D = m.D + static_cast<int64_t>(m.mat[e]) * 36;
  const double rho = m.rho[m.mat[e]];
  auto X = [&](int a, int c) { return m.nds[static_cast<int64_t>(v[a]) * m.dim + c]; };
  if (t == kCPE3 || t == kAC2D3) {
    const double x1 = X(0, 0), y1 = X(0, 1), x2 = X(1, 0), y2 = X(1, 1), x3 = X(2, 0), y3 = X(2, 1);
    const double J = x1 * y2 - y1 * x2 - x1 * y3 + y1 * x3 + x2 * y3 - y2 * x3;
    const double b[3] = {y2 - y3, y3 - y1, y1 - y2};
    const double c[3] = {x3 - x2, x1 - x3, x2 - x1};
    if (t == kAC2D3) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) K[i * 3 + j] = -(b[i] * b[j] + c[i] * c[j]) / (2.0 * J * rho);
      for (int i = 0; i < 3; ++i) mdiag[i] = -J / (6.0 * D[0]);
      return 3;
    }
    double B[6][9] = {};
    for (int i = 0; i < 3; ++i) {
      const double bi = b[i] / J, ci = c[i] / J;
      B[0][3 * i] = bi;
      B[1][3 * i + 1] = ci;
      B[3][3 * i + 2] = ci;
      B[4][3 * i + 2] = bi;
      B[5][3 * i] = ci;
      B[5][3 * i + 1] = bi;
    }
    double DB[6][9];
    for (int r = 0; r < 6; ++r)
      for (int j = 0; j < 9; ++j) {
        double s = 0.0;
        for (int q = 0; q < 6; ++q) s += D[r * 6 + q] * B[q][j];
        DB[r][j] = s;
      }
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j) {
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += B[r][i] * DB[r][j];
        K[i * 9 + j] = s * (J / 2.0);
      }
    for (int i = 0; i < 9; ++i) mdiag[i] = J * rho / 6.0;
    return 9;
  }
  // C3D4
  double E[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) E[r][c] = X(r + 1, c) - X(0, c);
  const double J = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) -
                   E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                   E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
  // inverse of E: Einv[j][k] = d xi_k / d x_j
  double inv[3][3];
  inv[0][0] = (E[1][1] * E[2][2] - E[1][2] * E[2][1]) / J;
  inv[0][1] = (E[0][2] * E[2][1] - E[0][1] * E[2][2]) / J;
  inv[0][2] = (E[0][1] * E[1][2] - E[0][2] * E[1][1]) / J;
  inv[1][0] = (E[1][2] * E[2][0] - E[1][0] * E[2][2]) / J;
  inv[1][1] = (E[0][0] * E[2][2] - E[0][2] * E[2][0]) / J;
  inv[1][2] = (E[0][2] * E[1][0] - E[0][0] * E[1][2]) / J;
  inv[2][0] = (E[1][0] * E[2][1] - E[1][1] * E[2][0]) / J;
  inv[2][1] = (E[0][1] * E[2][0] - E[0][0] * E[2][1]) / J;
  inv[2][2] = (E[0][0] * E[1][1] - E[0][1] * E[1][0]) / J;
  const double gref[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double G[4][3];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += gref[i][k] * inv[j][k];
      G[i][j] = s;
    }
  double B[6][12] = {};
  for (int i = 0; i < 4; ++i) {
    B[0][3 * i] = G[i][0];
    B[1][3 * i + 1] = G[i][1];
    B[2][3 * i + 2] = G[i][2];
    B[3][3 * i + 1] = G[i][2];
    B[3][3 * i + 2] = G[i][1];
    B[4][3 * i] = G[i][2];
    B[4][3 * i + 2] = G[i][0];
    B[5][3 * i] = G[i][1];
    B[5][3 * i + 1] = G[i][0];
  }
  double DB[6][12];
  for (int r = 0; r < 6; ++r)
    for (int j = 0; j < 12; ++j) {
      double s = 0.0;
      for (int q = 0; q < 6; ++q) s += D[r * 6 + q] * B[q][j];
      DB[r][j] = s;
    }
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) {
      double s = 0.0;
      for (int r = 0; r < 6; ++r) s += B[r][i] * DB[r][j];
      K[i * 12 + j] = s * (J / 6.0);
    }
  for (int i = 0; i < 12; ++i) mdiag[i] = J * rho / 24.0;
  return 12;
}

// full-DOF index of (local node a, DOF d) in element_matrix's layout
int full_index(int32_t t, int a, int d) { return t == kAC2D3 ? a : 3 * a + (d - 1); }

struct Work {
  Mesh m;
  int max_df = 0;
  std::vector<int64_t> adj_ptr;  // node -> elements
  std::vector<int64_t> adj;
  std::vector<int64_t> newidx;   // global (node*max_df + df-1) -> compacted index or -1
  int64_t n_dof = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<double> kval, mval, cval;
  std::vector<int64_t> gl;  // n_nodes x max_df, 1-based, 0 absent
};

// Row contributions of node v: (global index of column DOF -> value) for
// each DOF of v, plus the lumped mass and damping of v's DOFs.
void node_rows(const Work& W, int64_t v, std::vector<std::vector<std::pair<int64_t, double>>>& rows,
               std::vector<double>& mdiag, std::vector<double>& cdiag) {
  const Mesh& m = W.m;
  const int md = W.max_df;
  rows.assign(md, {});
  mdiag.assign(md, 0.0);
  cdiag.assign(md, 0.0);
  double K[144], Md[12];
  std::vector<int> ln, ld;
  for (int64_t q = W.adj_ptr[v]; q < W.adj_ptr[v + 1]; ++q) {
    const int64_t e = W.adj[q];
    const int32_t t = m.typ[e];
    element_matrix(m, e, K, Md);
    const int nf = t == kAC2D3 ? 3 : 3 * nodes_of(t);
    const double a = m.abs_i ? m.abs_i[e] : 0.0;
    const double damp = (a != 0.0) ? std::pow(a, m.p_law) * m.max_damp : 0.0;
    const double red = (a != 0.0) ? std::exp(m.log_msr * std::pow(a, m.p_law + 1.0)) : 1.0;
    local_dofs(t, m.dof_mask, ln, ld);
    const int32_t* ev = m.els + e * m.nper;
    for (size_t i = 0; i < ln.size(); ++i) {
      if (ev[ln[i]] != v) continue;
      const int fi = full_index(t, ln[i], ld[i]);
      const int r = ld[i] - 1;
      mdiag[r] += Md[fi];
      cdiag[r] += Md[fi] * damp;
      for (size_t j = 0; j < ln.size(); ++j) {
        const int fj = full_index(t, ln[j], ld[j]);
        const double kv = K[fi * nf + fj];
        if (kv == 0.0) continue;
        rows[r].emplace_back(static_cast<int64_t>(ev[ln[j]]) * md + (ld[j] - 1), kv * red);
      }
    }
  }
  for (auto& r : rows) {
    std::stable_sort(r.begin(), r.end(), [](auto& x, auto& y) { return x.first < y.first; });
    size_t o = 0;
    for (size_t i = 0; i < r.size();) {
      size_t j = i;
      double s = 0.0;
      while (j < r.size() && r[j].first == r[i].first) s += r[j++].second;
      if (s != 0.0) r[o++] = {r[i].first, s};
      i = j;
    }
    r.resize(o);
  }
}

thread_local std::string g_asm_error;

}  // namespace

extern "C" {

// Returns BFE_OK and the sizes; *work must be passed to bfe_assemble_fill.
bfe_status bfe_assemble_pattern(int64_t n_nodes, int32_t dim, const double* nds, int64_t n_el,
                                int32_t nper, const int32_t* els, const int32_t* el_type,
                                const int32_t* el_mat, const double* el_abs, int32_t n_mat,
                                const double* mat_D, const double* mat_rho, int32_t dof_mask,
                                double damping_power_law, double max_damping,
                                double max_stiffness_reduction, int64_t* n_dof, int64_t* nnz,
                                int32_t* max_df, void** work) {
  try {
    if (!n_dof || !nnz || !max_df || !work) throw std::invalid_argument("NULL output");
    *work = nullptr;
    if (n_nodes <= 0 || n_el <= 0 || (dim != 2 && dim != 3) || nper < 3)
      throw std::invalid_argument("bad mesh sizes");
    if (!nds || !els || !el_type || !el_mat || !mat_D || !mat_rho)
      throw std::invalid_argument("NULL mesh array");
    auto W = new Work();
    Mesh& m = W->m;
    m.n_nodes = n_nodes; m.dim = dim; m.nds = nds; m.n_el = n_el; m.nper = nper; m.els = els;
    m.typ = el_type; m.mat = el_mat; m.abs_i = el_abs; m.n_mat = n_mat; m.D = mat_D;
    m.rho = mat_rho; m.dof_mask = dof_mask; m.p_law = damping_power_law;
    m.max_damp = max_damping; m.log_msr = std::log(max_stiffness_reduction);
    // DOFs in use and max_df (fn_find_dof_in_use_and_max_dof_per_el.m)
    bool seen[3] = {false, false, false};
    int dmax = 0;
    std::vector<int> ln, ld;
    for (int64_t e = 0; e < n_el; ++e) {
      const int32_t t = el_type[e];
      if (t < 0 || t > 2) { delete W; throw std::invalid_argument("unsupported element type"); }
      if (el_mat[e] < 0 || el_mat[e] >= n_mat) { delete W; throw std::invalid_argument("material out of range"); }
      if (t == kC3D4 && dim != 3) { delete W; throw std::invalid_argument("C3D4 needs 3D nodes"); }
      if (!seen[t]) {
        seen[t] = true;
        local_dofs(t, dof_mask, ln, ld);
        for (int d : ld) dmax = std::max(dmax, d);
      }
      for (int a = 0; a < nodes_of(t); ++a) {
        const int32_t v = els[e * nper + a];
        if (v < 0 || v >= n_nodes) { delete W; throw std::invalid_argument("node index out of range"); }
      }
    }
    W->max_df = dmax;
    // node -> element adjacency
    W->adj_ptr.assign(n_nodes + 1, 0);
    for (int64_t e = 0; e < n_el; ++e)
      for (int a = 0; a < nodes_of(el_type[e]); ++a) W->adj_ptr[els[e * nper + a] + 1]++;
    for (int64_t v = 0; v < n_nodes; ++v) W->adj_ptr[v + 1] += W->adj_ptr[v];
    W->adj.resize(W->adj_ptr[n_nodes]);
    std::vector<int64_t> fill(W->adj_ptr.begin(), W->adj_ptr.end() - 1);
    for (int64_t e = 0; e < n_el; ++e)
      for (int a = 0; a < nodes_of(el_type[e]); ++a) W->adj[fill[els[e * nper + a]]++] = e;
    // pass 1: which DOFs have a non-empty column (K symmetric: row = column), row lengths
    const int md = W->max_df;
    std::vector<uint8_t> keep(static_cast<size_t>(n_nodes) * md, 0);
    std::vector<int64_t> len(static_cast<size_t>(n_nodes) * md, 0);
    par_for(n_nodes, [&](int64_t v) {
      std::vector<std::vector<std::pair<int64_t, double>>> rows;
      std::vector<double> mdg, cdg;
      node_rows(*W, v, rows, mdg, cdg);
      for (int d = 0; d < md; ++d) {
        keep[v * md + d] = (!rows[d].empty() || mdg[d] != 0.0 || cdg[d] != 0.0) ? 1 : 0;
        len[v * md + d] = static_cast<int64_t>(rows[d].size());
      }
    });
    W->newidx.assign(static_cast<size_t>(n_nodes) * md, -1);
    int64_t k = 0;
    W->rowptr.push_back(0);
    for (int64_t g = 0; g < n_nodes * md; ++g)
      if (keep[g]) {
        W->newidx[g] = k++;
        W->rowptr.push_back(W->rowptr.back() + len[g]);
      }
    W->n_dof = k;
    if (k > INT32_MAX) { delete W; throw std::invalid_argument("more than 2^31 DOF"); }
    *n_dof = k;
    *nnz = W->rowptr.back();
    *max_df = md;
    *work = W;
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    g_asm_error = e.what();
    return BFE_ERR_ARG;
  } catch (const std::bad_alloc&) {
    g_asm_error = "out of memory";
    return BFE_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_asm_error = e.what();
    return BFE_ERR_INTERNAL;
  }
}

// Fills the CSR of K (rowptr n_dof+1, col/val nnz), the lumped mass m and
// the damping diagonal c (n_dof each), and gl_lookup (n_nodes x max_df,
// row-major, 1-based, 0 = absent); releases the work object.
bfe_status bfe_assemble_fill(void* work, int64_t* rowptr, int32_t* col, double* kval, double* m,
                             double* c, int64_t* gl_lookup) {
  Work* W = static_cast<Work*>(work);
  try {
    if (!W) throw std::invalid_argument("NULL work");
    if (!rowptr || !col || !kval || !m || !c || !gl_lookup) throw std::invalid_argument("NULL output");
    const int md = W->max_df;
    const int64_t nn = W->m.n_nodes;
    std::memcpy(rowptr, W->rowptr.data(), sizeof(int64_t) * (W->n_dof + 1));
    par_for(nn, [&](int64_t v) {
      std::vector<std::vector<std::pair<int64_t, double>>> rows;
      std::vector<double> mdg, cdg;
      node_rows(*W, v, rows, mdg, cdg);
      for (int d = 0; d < md; ++d) {
        const int64_t g = v * md + d;
        gl_lookup[g] = W->newidx[g] >= 0 ? W->newidx[g] + 1 : 0;
        const int64_t r = W->newidx[g];
        if (r < 0) continue;
        m[r] = mdg[d];
        c[r] = cdg[d];
        int64_t o = W->rowptr[r];
        for (auto& en : rows[d]) {
          const int64_t cj = W->newidx[en.first];
          if (cj < 0) throw std::runtime_error("non-symmetric element stiffness pattern");
          col[o] = static_cast<int32_t>(cj);
          kval[o] = en.second;
          ++o;
        }
      }
    });
    delete W;
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    g_asm_error = e.what();
    delete W;
    return BFE_ERR_ARG;
  } catch (const std::exception& e) {
    g_asm_error = e.what();
    delete W;
    return BFE_ERR_INTERNAL;
  }
}

const char* bfe_assemble_last_error(void) { return g_asm_error.c_str(); }

}  // extern "C"

namespace bfe {
// The device assembler (bfe_assemble_dev.hip) reports through the same string.
void set_assembly_error(const std::string& msg) { g_asm_error = msg; }
}  // namespace bfe

// Host-side FP64 set-up of the explicit march (no device code here).
//
// Follows code/internal/fn_explicit_dynamic_solver_v6.m:122-139: the lumped
// mass m = sum(M), and the update u = A f + B u1 + D u2 with
//   'vel at curent time step':  X = I + dt/2 M^-1 C,
//        A = X \ dt^2 M^-1,  B = X \ (2I - dt^2 M^-1 K),  D = X \ (-I + dt/2 M^-1 C)
//   'vel at last half time step':
//        A = dt^2 M^-1,  B = 2I - dt M^-1 C - dt^2 M^-1 K,  D = dt M^-1 C - I
// Instead of forming A, B, D, every row keeps
//   u_i = beta_i u1_i + gamma_i u2_i - sum_j (alpha_i K_ij) u1_j + alpha_i f_i
// which is exact for rows where C has no off-diagonal entries. Rows coupled by
// off-diagonal C (ASI2D2 fluid-solid interface, fn_el_ASI2D2.m:60-76) form the
// set I; they get an explicit correction and, in the default mode, the exact
// solve u_I = X_II^-1 y_I with X_II^-1 computed here in FP64.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bfe {

struct Csr {
  int64_t n = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> col;
  std::vector<double> val;
  int64_t nnz() const { return ptr.empty() ? 0 : ptr.back(); }
};

// Read-only view of a CSR held elsewhere (the caller's arrays, or a Csr).
struct CsrRef {
  int64_t n = 0;
  const int64_t* ptr = nullptr;
  const int32_t* col = nullptr;
  const double* val = nullptr;
  int64_t nnz() const { return ptr ? ptr[n] : 0; }
};
inline CsrRef ref(const Csr& c) { return {c.n, c.ptr.data(), c.col.data(), c.val.data()}; }

// Sliced ELLPACK with slice height 64 (one wavefront), no row sorting.
// Entry k of row r lives at off[r/64] + 64*k + r%64. Padding entries point
// at the row itself with value 0.
struct Sell64 {
  int64_t n = 0;
  int64_t n_slices = 0;
  std::vector<int64_t> off;  // n_slices + 1, in entries
  std::vector<int32_t> col;
  std::vector<double> val;
};

struct Plan {
  int64_t n = 0;
  int mode = 0;
  double dt = 0.0;
  std::vector<double> alpha, beta, gamma;
  bool damped = false;  // some beta_i != 2 (gamma_i = 1 - beta_i is derived)
  int64_t nnz_k = 0;
  Sell64 op;  // alpha-scaled K
  // Interface rows (sorted) and their off-diagonal C scaled by 1/m_i.
  std::vector<int32_t> iface;         // global row of interface entry k
  std::vector<int32_t> iface_pos;     // n entries: k or -1
  Csr coff;                           // rows k over global columns
  double sq = 0.0, sr = 0.0;          // y += sq*Coff*u1 + sr*Coff*u2
  Csr xinv;                           // rows/cols over interface positions
  bool has_xinv = false;
  // gamma_i = 1 - beta_i on every non-interface row (both modes) and on
  // interface rows of the last-half-step mode; gfix_k = gamma - (1 - beta)
  // on interface row k (dt c_i / (2 m_i) in the current-step mode).
  std::vector<double> gfix;
};

// Normalises a CSR (sorted columns, duplicates summed, exact zeros removed)
// and validates it. Throws std::invalid_argument on malformed input.
Csr normalise_csr(int64_t n, const int64_t* ptr, const int32_t* col,
                  const double* val, const char* name);

// True when the CSR is already normal: row pointers non-decreasing from 0,
// columns in range and strictly increasing within each row, every value
// finite and non-zero (checked in parallel). Such a CSR is used in place.
bool csr_is_normal(int64_t n, const int64_t* ptr, const int32_t* col, const double* val);

// Builds the whole plan; with host_sell = false the SELL-64 operator is left
// empty (the caller builds it on the device from K and alpha). Throws
// std::invalid_argument (bad input) or std::domain_error (numerical
// problem, e.g. m_i == 0).
Plan build_plan(int64_t n, const CsrRef& K, const double* m, const Csr& C,
                double dt, int mode, bool host_sell = true);

// Exact inverse of the sparse, diagonally dominated interface block, column
// by column on growing graph neighbourhoods (dense LU with partial pivoting
// on each neighbourhood) until the neighbourhood boundary carries less than
// 1e-20 of the column; entries below drop_tol of a column's largest entry
// are dropped. With weights w (empty: none) the columns are solved on the
// equilibrated block W X W^-1 (W = diag(w)) and both tests judge its
// inverse: with X = I + dt/2 M^-1 C and w = sqrt|m| that block is
// I + dt/2 |M|^-1/2 sign(M) C |M|^-1/2, whose condition number on the
// fluid-solid interface is about 1 against 1.7e10 for X (|m| spans 13
// orders of magnitude; without the weights the histories moved by up to
// 1.5e-10 through dropped entries and by 5e-13 through the solves).
Csr truncated_inverse(const Csr& X, double drop_tol, const std::vector<double>& w = {});

}  // namespace bfe

// MI355X (gfx950) kernels and C ABI of the explicit time-march.
//
// One time step of code/internal/fn_explicit_dynamic_solver_v6.m:180-212
//   f(idx) = F(:, ti);  u = A f + B u1 + D u2;  hist(:, ti) = u(h_idx);
//   [fld(:, fi) = (u - u1)/dt];  u2 = u1;  u1 = u
// becomes one fused, HBM-streaming kernel over all rows (k_step) plus, only
// when the model has fluid-solid interface rows in the current-step mode, one
// small kernel (k_iface) that applies the implicit interface block
// u_I = X_II^-1 y_I to the right-hand side y_I that k_step left in a compact
// buffer.
// Monitor gathers and field frames of step t-1 are written by the step-t
// kernel from the u1/u2 values it already reads, so they cost no extra pass.
//
// Operator layout (host set-up in bfe_plan.cpp): SELL-64 of alpha-scaled K.
// A wavefront owns one 64-row slice; lane l streams row 64*s+l, so the value
// and column streams are read fully coalesced (64 consecutive entries per
// load) and no cross-lane reduction is needed.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bfe.h"
#include "bfe_ebe.hpp"
#include "bfe_plan.hpp"

namespace {

constexpr int kBlock = 256;              // threads per workgroup
constexpr int kRowsSingle = kBlock;      // rows per workgroup, single source

struct hip_failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define BFE_HIP(expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      throw hip_failure(std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

thread_local std::string g_last_error;

// Logical block index: consecutive logical blocks stay on one XCD, so the
// u1 gathers of neighbouring rows hit that XCD's L2. Hardware dispatches
// workgroups round-robin over the 8 XCDs; the mapping is a bijection for any
// grid size, so correctness never depends on the dispatch order.
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nb) {
  const int64_t q = nb >> 3, r = nb & 7;
  const int64_t x = bid & 7, j = bid >> 3;
  return x * q + (x < r ? x : r) + j;
}

// Per-workgroup lists (forcing, monitors, interface rows) let a thread find
// the few special rows of its block without a per-row array in HBM.
struct BlockLists {
  const int32_t* fptr;    // forcing: block b owns [fptr[b], fptr[b+1])
  const int32_t* f_row;
  const int32_t* f_src;   // source column (multi-source)
  const int32_t* f_frow;  // row of the signal matrix F
  const int32_t* mptr;    // monitors
  const int32_t* m_row;
  const int32_t* m_k;
  const int32_t* iptr;    // interface rows
  const int32_t* i_row;
  const int32_t* i_k;
};

// Interface rows k (global row iface[k]): off-diagonal C scaled by 1/m,
// the gamma correction and the solver-mode factors. k_step turns w into
//   y_k = w_k + gfix_k u2_r + sq sum_j Coff_kj u1_j + sr sum_j Coff_kj u2_j
// which is final in the last-half-step mode and the right-hand side of
// X_II u_I = y_I (applied by k_iface) in the current-step mode.
template <typename T>
struct IfaceRows {
  const int64_t* cptr;
  const int32_t* ccol;
  const T* cval;
  const T* gfix;
  T sq, sr;
  T* wI;  // nullptr: no X_II solve, y is final
};

// ---------------------------------------------------------------------------
// Row-per-thread step over S columns (S = 1: the march; S = 2: the adjoint
// sweep that carries a recomputed forward column and the adjoint column
// through one read of the operator). u is row-major n x S.
//   u[r,s] = beta[r] u1[r,s] + (1 - beta[r]) u2[r,s] - sum_k val[k] u1[col[k],s] (+ forcing)
// (gamma = 1 - beta holds on every row outside the interface; interface rows
// add their correction from IfaceRows). DAMPED = false: beta = 2 everywhere
// and no beta array is read. store.p[s], when set, receives column s of u.
template <typename T, int S>
struct ColPtrs {
  T* p[S];
};

// Load and store of a stream read or written once per step: with NT the
// access is non-temporal, so it does not push the neighbours' u1 lines out
// of the caches.
template <bool NT, typename T>
__device__ __forceinline__ T stream_load(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void stream_store(T v, T* p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Operators larger than the 256 MB Infinity Cache stream their step-only
// arrays non-temporally (measured: CSR step 248 -> 227 us on the 7.7M-DOF
// coupled model; 2-4 % slower on a 40k-DOF plate that fits in cache).
constexpr int64_t kStreamBytes = 256ll << 20;

// One row of workgroup block lb (the body of k_step and of the persistent
// march).
template <typename T, int S, bool DAMPED, bool NT>
__device__ __forceinline__ void step_row(
    int64_t lb, int64_t row, int64_t n, const int64_t* __restrict__ off,
    const int32_t* __restrict__ col, const T* __restrict__ val, const T* __restrict__ beta,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u,
    const BlockLists& L, const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows,
    int32_t t, int32_t q, T* __restrict__ hist_prev, T* __restrict__ fld_prev,
    int32_t fld_is_vel, T dt, const IfaceRows<T>& I, const ColPtrs<T, S>& store) {
  if (row >= n) return;
  const int64_t sl = row >> 6;
  const int lane = static_cast<int>(row & 63);
  const int64_t e = off[sl + 1];
  int64_t p = off[sl] + lane;
  T acc[S], acc2[S];
#pragma unroll
  for (int s = 0; s < S; ++s) acc[s] = acc2[s] = T(0);
  for (; p + 64 < e; p += 128) {
    const T v0 = stream_load<NT>(val + p), v1 = stream_load<NT>(val + p + 64);
    const int64_t c0 = static_cast<int64_t>(stream_load<NT>(col + p)) * S,
                  c1 = static_cast<int64_t>(stream_load<NT>(col + p + 64)) * S;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      acc[s] += v0 * u1[c0 + s];
      acc2[s] += v1 * u1[c1 + s];
    }
  }
  if (p < e) {
    const T v0 = stream_load<NT>(val + p);
    const int64_t c0 = static_cast<int64_t>(stream_load<NT>(col + p)) * S;
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] += v0 * u1[c0 + s];
  }
  T a[S], b[S], un[S];
  const T be = DAMPED ? stream_load<NT>(beta + row) : T(2);
#pragma unroll
  for (int s = 0; s < S; ++s) {
    a[s] = u1[row * S + s];
    b[s] = stream_load<NT>(u2 + row * S + s);
    if constexpr (DAMPED) un[s] = be * a[s] + (T(1) - be) * b[s] - (acc[s] + acc2[s]);
    else un[s] = (T(2) * a[s] - b[s]) - (acc[s] + acc2[s]);
  }
  for (int32_t i = L.fptr[lb], ie = L.fptr[lb + 1]; i < ie; ++i)
    if (L.f_row[i] == row) {
      const T f = f_alpha[i] * F[static_cast<int64_t>(t) * F_rows + L.f_frow[i]];
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (L.f_src[i] == s) un[s] += f;
    }
  for (int32_t i = L.iptr[lb], ie = L.iptr[lb + 1]; i < ie; ++i)
    if (L.i_row[i] == row) {
      const int32_t k = L.i_k[i];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        T c1 = T(0), c2 = T(0);
        for (int64_t j = I.cptr[k], je = I.cptr[k + 1]; j < je; ++j) {
          const int64_t c = static_cast<int64_t>(I.ccol[j]) * S + s;
          c1 += I.cval[j] * u1[c];
          c2 += I.cval[j] * u2[c];
        }
        un[s] += I.gfix[k] * b[s] + I.sq * c1 + I.sr * c2;
        if (I.wI != nullptr) I.wI[static_cast<int64_t>(k) * S + s] = un[s];
      }
    }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    stream_store<NT>(un[s], u + row * S + s);
    if (store.p[s] != nullptr) store.p[s][row] = un[s];
  }
  if constexpr (S == 1)
    if (fld_prev != nullptr) fld_prev[row] = fld_is_vel ? (a[0] - b[0]) / dt : a[0];
  if (hist_prev != nullptr)
    for (int32_t i = L.mptr[lb], ie = L.mptr[lb + 1]; i < ie; ++i)
      if (L.m_row[i] == row) {
#pragma unroll
        for (int s = 0; s < S; ++s) hist_prev[L.m_k[i] + static_cast<int64_t>(q) * s] = a[s];
      }
}

template <typename T, int S, bool DAMPED, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_step(
    int64_t n, int64_t nblk, const int64_t* __restrict__ off,
    const int32_t* __restrict__ col, const T* __restrict__ val, const T* __restrict__ beta,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u,
    BlockLists L, const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows,
    int32_t t, int32_t q, T* __restrict__ hist_prev, T* __restrict__ fld_prev,
    int32_t fld_is_vel, T dt, IfaceRows<T> I, ColPtrs<T, S> store) {
  const int64_t lb = xcd_block(blockIdx.x, nblk);
  step_row<T, S, DAMPED, NT>(lb, lb * kRowsSingle + threadIdx.x, n, off, col, val, beta, u1, u2,
                             u, L, f_alpha, F, F_rows, t, q, hist_prev, fld_prev, fld_is_vel, dt,
                             I, store);
}

// ---------------------------------------------------------------------------
// Multi right-hand side step (S <= 64 sources, u stored row-major n x S).
// One wavefront per row, lane = source: the row's K entries are wave-uniform
// and the u1 gathers read S contiguous values. A workgroup holds blockDim/64
// rows (tunable); the per-row lists (forcing, monitors, interface) are indexed
// by row. remap = 1 applies the XCD-aware workgroup order.
// R consecutive rows of one SELL-64 slice per wavefront (R divides 64, so they
// share the slice width): their entry loads, shuffles and u1 gathers are
// interleaved, R independent chains in flight per lane. The forcing,
// interface and monitor lists are grouped per workgroup (lb).
// G (rows per gather group): with G > 1 the operator stores each aligned
// group of G rows over the union of their columns (explicit zeros where a
// row lacks a column; G = 2 pairs the two displacement rows of a 2D solid
// node, G = 4 two neighbouring nodes), so one u1 gather serves G rows and
// only the first row's column indices are loaded. The zero terms only move
// the two-term grouping of a row's sum: results agree to rounding.
// Entries are broadcast from the lane that loaded them with readlane (VALU,
// into scalar registers), not through the LDS crossbar.
__device__ __forceinline__ int lane_bcast(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ __forceinline__ float lane_bcast(float v, int k) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
}
__device__ __forceinline__ double lane_bcast(double v, int k) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), k);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), k);
  return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) |
                                        static_cast<long long>(static_cast<unsigned>(lo)));
}

template <typename T, bool DAMPED, int R, int G>
__device__ __forceinline__ void multi_rows(
    int64_t lb, int64_t row0, int64_t n, int32_t S, int lane, const int64_t* __restrict__ off,
    const int32_t* __restrict__ col, const T* __restrict__ val, const T* __restrict__ beta,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u, const BlockLists& L,
    const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows, int32_t t, int32_t q,
    T* __restrict__ hist_prev, const IfaceRows<T>& I) {
  const bool active = lane < S;
  const int64_t sl = row0 >> 6;
  const int w = static_cast<int>((off[sl + 1] - off[sl]) >> 6);  // entries per row of the slice
  const int64_t b0 = off[sl] + (row0 & 63) * static_cast<int64_t>(w);  // rows contiguous
  // The rows' entries are loaded once, one per lane, and broadcast with
  // cross-lane shuffles: the gathers of u1 then issue back to back.
  static_assert(G >= 1 && R % G == 0, "rows per wavefront must be a multiple of the group");
  constexpr int RS = G;  // rows served by one gather
  constexpr int NG = R / G;  // row groups per wavefront
  constexpr int U = NG >= 4 ? 2 : NG == 2 ? 4 : 8;
  T acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = T(0);
  for (int k0 = 0; k0 < w; k0 += 64) {
    const int cnt = min(64, w - k0);
    T myv[R];
    int myc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      myv[r] = T(0);
      myc[r] = 0;
      if (lane < cnt) {
        myv[r] = val[b0 + static_cast<int64_t>(r) * w + k0 + lane];
        if (r % RS == 0) myc[r] = col[b0 + static_cast<int64_t>(r) * w + k0 + lane];
      }
    }
    int k = 0;
    // U gathers per row group issued back to back (about 8 per wavefront in
    // flight: the step is bound by gather latency), summed in the same
    // pairs, in the same order, as the two-entry loop below
    for (; k + U <= cnt; k += U) {
#pragma unroll
      for (int r = 0; r < R; r += RS) {
        T x[U];
#pragma unroll
        for (int e = 0; e < U; ++e) {
          const int c = lane_bcast(myc[r], k + e);
          x[e] = T(0);
          if (active) x[e] = u1[static_cast<int64_t>(c) * S + lane];
        }
#pragma unroll
        for (int j = 0; j < RS; ++j)
#pragma unroll
          for (int e = 0; e < U; e += 2) {
            const T v0 = lane_bcast(myv[r + j], k + e), v1 = lane_bcast(myv[r + j], k + e + 1);
            acc[r + j] += v0 * x[e] + v1 * x[e + 1];
          }
      }
    }
    for (; k + 2 <= cnt; k += 2) {
#pragma unroll
      for (int r = 0; r < R; r += RS) {
        const int c0 = lane_bcast(myc[r], k), c1 = lane_bcast(myc[r], k + 1);
        T x0 = T(0), x1 = T(0);
        if (active) {
          x0 = u1[static_cast<int64_t>(c0) * S + lane];
          x1 = u1[static_cast<int64_t>(c1) * S + lane];
        }
#pragma unroll
        for (int j = 0; j < RS; ++j) {
          const T v0 = lane_bcast(myv[r + j], k), v1 = lane_bcast(myv[r + j], k + 1);
          acc[r + j] += v0 * x0 + v1 * x1;
        }
      }
    }
    for (; k < cnt; ++k) {
#pragma unroll
      for (int r = 0; r < R; r += RS) {
        const int c0 = lane_bcast(myc[r], k);
        T x0 = T(0);
        if (active) x0 = u1[static_cast<int64_t>(c0) * S + lane];
#pragma unroll
        for (int j = 0; j < RS; ++j) {
          const T v0 = lane_bcast(myv[r + j], k);
          if (active) acc[r + j] += v0 * x0;
        }
      }
    }
  }
  if (!active) return;
  const int nr = static_cast<int>(min<int64_t>(R, n - row0));
  T a[R], b[R], un[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t row = r < nr ? row0 + r : row0;  // rows past n recompute row0, then drop it
    const int64_t me = row * S + lane;
    a[r] = u1[me];
    // u2 and u are n x S streams (2.5 GB each at 4.9M DOF x 64 sources):
    // non-temporal, so they leave the caches to the u1 gathers (2 % faster
    // on that workload; streaming the operator too made it 7 % slower)
    b[r] = stream_load<true>(u2 + me);
    if constexpr (DAMPED) {
      const T be = beta[row];
      un[r] = be * a[r] + (T(1) - be) * b[r] - acc[r];
    } else {
      un[r] = (T(2) * a[r] - b[r]) - acc[r];
    }
  }
  // forcing, interface and monitor entries of the workgroup (usually none)
  for (int32_t i = L.fptr[lb], ie = L.fptr[lb + 1]; i < ie; ++i) {
    const int64_t rr = L.f_row[i] - row0;
    if (rr >= 0 && rr < nr && L.f_src[i] == lane) {
      const T f = f_alpha[i] * F[static_cast<int64_t>(t) * F_rows + L.f_frow[i]];
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (rr == r) un[r] += f;
    }
  }
  for (int32_t i = L.iptr[lb], ie = L.iptr[lb + 1]; i < ie; ++i) {
    const int64_t rr = L.i_row[i] - row0;
    if (rr < 0 || rr >= nr) continue;
    const int32_t k = L.i_k[i];
    T c1 = T(0), c2 = T(0);
    for (int64_t p = I.cptr[k], pe = I.cptr[k + 1]; p < pe; ++p) {
      const int64_t c = static_cast<int64_t>(I.ccol[p]) * S + lane;
      c1 += I.cval[p] * u1[c];
      c2 += I.cval[p] * u2[c];
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (rr == r) {
        un[r] += I.gfix[k] * b[r] + I.sq * c1 + I.sr * c2;
        if (I.wI != nullptr) I.wI[static_cast<int64_t>(k) * S + lane] = un[r];
      }
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (r < nr) stream_store<true>(un[r], u + (row0 + r) * S + lane);
  if (hist_prev != nullptr)
    for (int32_t i = L.mptr[lb], ie = L.mptr[lb + 1]; i < ie; ++i) {
      const int64_t rr = L.m_row[i] - row0;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (rr == r && r < nr) hist_prev[L.m_k[i] + static_cast<int64_t>(q) * lane] = a[r];
    }
}

// Wavefront v of the grid marches rows v*R .. v*R + R - 1.
template <typename T, bool DAMPED, int R, int G>
__global__ __launch_bounds__(512) void k_step_multi(
    int64_t n, int32_t S, int64_t nblk, int32_t remap, const int64_t* __restrict__ off,
    const int32_t* __restrict__ col, const T* __restrict__ val, const T* __restrict__ beta,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u, BlockLists L,
    const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows, int32_t t, int32_t q,
    T* __restrict__ hist_prev, IfaceRows<T> I) {
  const int64_t lb = remap ? xcd_block(blockIdx.x, nblk) : static_cast<int64_t>(blockIdx.x);
  // wavefront index as a scalar: row and entry offsets stay in SGPRs
  const int64_t wave =
      lb * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int64_t row0 = wave * R;
  if (row0 >= n) return;  // uniform across the wavefront
  multi_rows<T, DAMPED, R, G>(lb, row0, n, S, lane, off, col, val, beta, u1, u2, u, L, f_alpha, F,
                           F_rows, t, q, hist_prev, I);
}

// ---------------------------------------------------------------------------
// Interface solve of the current-step mode: u[iface[k]] = sum_l Xinv_kl y_l
// with y = wI from k_step (compact, L2-resident). Single source: 16 lanes
// share one row and reduce with cross-lane shuffles, so each lane walks only
// a few of the row's entries. Reads wI, writes interface rows of u only.
constexpr int kIfaceLanes = 16;

template <typename T>
__device__ __forceinline__ void iface_lane(
    int64_t gid, int32_t nI, const int32_t* __restrict__ iface, const int64_t* __restrict__ xptr,
    const int32_t* __restrict__ xcol, const T* __restrict__ xval, const T* __restrict__ wI,
    T* __restrict__ u) {
  const int64_t k = gid / kIfaceLanes;
  const int j = static_cast<int>(gid % kIfaceLanes);
  T acc = T(0);
  if (k < nI)
    for (int64_t p = xptr[k] + j, pe = xptr[k + 1]; p < pe; p += kIfaceLanes)
      acc += xval[p] * wI[xcol[p]];
  for (int o = kIfaceLanes / 2; o > 0; o >>= 1) acc += __shfl_down(acc, o, kIfaceLanes);
  if (k < nI && j == 0) u[iface[k]] = acc;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_iface(
    int32_t nI, const int32_t* __restrict__ iface, const int64_t* __restrict__ xptr,
    const int32_t* __restrict__ xcol, const T* __restrict__ xval, const T* __restrict__ wI,
    T* __restrict__ u) {
  iface_lane<T>(static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x, nI, iface, xptr,
                xcol, xval, wI, u);
}

// ---------------------------------------------------------------------------
// Persistent multi-step march: one cooperative launch runs the whole time
// loop of a single-source CSR march, for meshes small enough that a step is
// bound by launch latency rather than by HBM. Every step the workgroups
// sweep the blocks of k_step (workgroup g takes blocks g, g + grid, ..., so
// the per-workgroup forcing, monitor and interface lists are those of
// k_step), meet at a grid barrier, and meet again after the interface solve
// when the plan has one. The loop bounds are kernel arguments, the same for
// every workgroup, so every workgroup reaches every barrier.
template <typename T>
struct PersistentMarch {
  int64_t n, nblk, t0, n_t;
  const int64_t* off;
  const int32_t* col;
  const T* val;
  const T* beta;
  T* ubuf;  // 3 n: u, u1, u2 of the first step, rotated as in the host loop
  BlockLists L;
  const T* f_alpha;
  const T* F;
  int32_t F_rows, q;
  T* hist;
  int32_t field_every;  // 0: no field frames
  int32_t fld_is_vel;
  T* fld;
  T dt;
  IfaceRows<T> I;
  int32_t nI;  // interface rows solved after each step (0: none)
  const int32_t* iface;
  const int64_t* xptr;
  const int32_t* xcol;
  const T* xval;
};

template <typename T, bool DAMPED, bool NT>
__global__ __launch_bounds__(kBlock) void k_march_persistent(PersistentMarch<T> a) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const ColPtrs<T, 1> store{{nullptr}};
  for (int64_t t = a.t0; t < a.n_t; ++t) {
    // the host loop's rotation (u2, u1, u) <- (u1, u, u2) after each step
    const int64_t k = (t - a.t0) % 3;
    T* u = a.ubuf + ((3 - k) % 3) * a.n;
    const T* u1 = a.ubuf + ((4 - k) % 3) * a.n;
    const T* u2 = a.ubuf + ((5 - k) % 3) * a.n;
    const int64_t prev = t - 1;
    T* hp = (a.q > 0 && prev >= a.t0) ? a.hist + prev * a.q : nullptr;
    T* fp = (a.field_every > 0 && prev >= a.t0 && prev % a.field_every == 0)
                ? a.fld + (prev / a.field_every) * a.n
                : nullptr;
    for (int64_t lb = blockIdx.x; lb < a.nblk; lb += gridDim.x)
      step_row<T, 1, DAMPED, NT>(lb, lb * kRowsSingle + threadIdx.x, a.n, a.off, a.col, a.val,
                                 a.beta, u1, u2, u, a.L, a.f_alpha, a.F, a.F_rows,
                                 static_cast<int32_t>(t), a.q, hp, fp, a.fld_is_vel, a.dt, a.I,
                                 store);
    grid.sync();
    if (a.nI > 0) {
      const int64_t work = static_cast<int64_t>(a.nI) * kIfaceLanes;
      for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < work;
           base += static_cast<int64_t>(gridDim.x) * blockDim.x)
        iface_lane<T>(base + threadIdx.x, a.nI, a.iface, a.xptr, a.xcol, a.xval, a.I.wI, u);
      grid.sync();
    }
  }
}

// Multi-source: one thread per (k, s); wI and u rows are S-contiguous.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_iface_multi(
    int32_t nI, int32_t S, const int32_t* __restrict__ iface, const int64_t* __restrict__ xptr,
    const int32_t* __restrict__ xcol, const T* __restrict__ xval, const T* __restrict__ wI,
    T* __restrict__ u) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(nI) * S) return;
  const int64_t k = idx / S, s = idx - k * S;
  T acc = T(0);
  for (int64_t p = xptr[k], pe = xptr[k + 1]; p < pe; ++p)
    acc += xval[p] * wI[static_cast<int64_t>(xcol[p]) * S + s];
  u[static_cast<int64_t>(iface[k]) * S + s] = acc;
}

// Monitors (and, single-source, the field frame) of the last step.
template <typename T>
__global__ void k_snap(int64_t n, int32_t S, const T* __restrict__ u_last,
                       const T* __restrict__ u_prev, int32_t q,
                       const int32_t* __restrict__ hist_idx, T* __restrict__ hist_last,
                       T* __restrict__ fld, int32_t fld_is_vel, T dt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (fld != nullptr && i < n) {
    const T a = u_last[i];
    fld[i] = fld_is_vel ? (a - u_prev[i]) / dt : a;
  }
  if (hist_last != nullptr && i < static_cast<int64_t>(q) * S) {
    const int64_t k = i % q, s = i / q;
    hist_last[k + q * s] = u_last[static_cast<int64_t>(hist_idx[k]) * S + s];
  }
}

// ---------------------------------------------------------------------------
// Element-by-element step (single source). Within one element group the
// stiffness is held as
//   K = sum_e s_e P_e' T[tau_e] P_e
// with a few element templates T (one per element shape and material, local
// DOF order node-major) and a per-element scale s_e (the absorbing-layer
// stiffness reduction times J_rep / J_e), and the global DOFs of the group's
// node v are r0 + c, r0 = dof0[v] (v*DPN when there is no DOF table). Thread v
// owns node v's DPN rows and gathers its row of K u from the elements around
// it: slot k of every node holds the element in which the node is local
// vertex l_k of template tau_k (or -1), so the DPN x (NPE*DPN) template rows a
// slot needs are the same for every lane and are read through the scalar
// cache. The slot-major adjacency is read coalesced; element records (each
// vertex's first DOF) and u1 are gathered. The update, forcing, interface,
// monitor and field handling are those of k_step on the node's rows.
constexpr int kEbeSlotPad = 8;  // slot count padding: any load group G of 1, 2, 4, 8 divides it

// acc[c] += sum over the G slots k0..k0+G-1 of s_e T_k(c, :) u1(e), given
// the slots' table entries e0 and element records nd, s.
template <typename T, int NPE, int DPN, int G>
__device__ __forceinline__ void ebe_group_apply(int32_t k0, const int32_t* e0, const int4* nd,
                                                const T* s, const T* __restrict__ srows,
                                                const T* __restrict__ u1, T* acc) {
  constexpr int KE = NPE * DPN;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const T* R = srows + static_cast<int64_t>(k0 + g) * (DPN * KE);
    const int32_t w[4] = {nd[g].x, nd[g].y, nd[g].z, nd[g].w};
    T y[DPN];
#pragma unroll
    for (int c = 0; c < DPN; ++c) y[c] = T(0);
#pragma unroll
    for (int b = 0; b < NPE; ++b) {
#pragma unroll
      for (int d = 0; d < DPN; ++d) {
        const T x = u1[static_cast<int64_t>(w[b]) + d];  // w: the vertex's first DOF
#pragma unroll
        for (int c = 0; c < DPN; ++c) y[c] += R[c * KE + b * DPN + d] * x;
      }
    }
#pragma unroll
    for (int c = 0; c < DPN; ++c) acc[c] += e0[g] < 0 ? T(0) : s[g] * y[c];
  }
}

template <typename T, int G>
__device__ __forceinline__ void ebe_load_table(int64_t n_nodes, int64_t v, int32_t k0,
                                               const int32_t* __restrict__ adj, int32_t* e0) {
#pragma unroll
  for (int g = 0; g < G; ++g) e0[g] = adj[static_cast<int64_t>(k0 + g) * n_nodes + v];
}

template <typename T, int G>
__device__ __forceinline__ void ebe_load_records(const int32_t* e0, const int4* __restrict__ el_nodes,
                                                 const T* __restrict__ el_scale, int4* nd, T* s) {
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int32_t e = e0[g] < 0 ? 0 : e0[g];  // an empty slot reads element 0, then drops it
    nd[g] = el_nodes[e];
    s[g] = el_scale[e];
  }
}

// a if bit == 1, b if bit == 0, by bit masking (no lane-mask predicate).
__device__ __forceinline__ double pick_bits(double a, double b, uint32_t bit) {
  const uint64_t m = 0ull - static_cast<uint64_t>(bit);
  return __longlong_as_double(static_cast<long long>(
      (static_cast<uint64_t>(__double_as_longlong(a)) & m) |
      (static_cast<uint64_t>(__double_as_longlong(b)) & ~m)));
}
__device__ __forceinline__ float pick_bits(float a, float b, uint32_t bit) {
  const uint32_t m = 0u - bit;
  return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}

// The update of node v's DPN rows (global rows r0 .. r0 + DPN - 1) from
// acc = (K u1) on those rows, with the forcing, interface, field and monitor
// handling of k_step; the lists hold the group's local rows v*DPN + c. alpha
// and beta are per DOF, or per node when coef_per_node is set. NT: u2,
// alpha and u are non-temporal streams.
template <typename T, bool DAMPED, int DPN, bool NT = false>
__device__ __forceinline__ void ebe_finish(
    int64_t v, int64_t lb, int64_t r0, const T* acc, const T* __restrict__ alpha,
    const T* __restrict__ beta, const T* __restrict__ u1, const T* __restrict__ u2,
    T* __restrict__ u, const BlockLists& L, const T* __restrict__ f_alpha,
    const T* __restrict__ F, int32_t F_rows, int32_t t, int32_t q, T* __restrict__ hist_prev,
    T* __restrict__ fld_prev, int32_t fld_is_vel, T dt, const IfaceRows<T>& I,
    int32_t coef_per_node, uint32_t undamped = 0u) {
  const int64_t l0 = v * DPN;  // the node's rows in the group's local row numbering
  // undamped nodes (every DOF with beta = 2 exactly, the rows outside the
  // absorbing layers) read beta entry 0 instead of their own: one cached
  // line per wave instead of the beta stream
  T a[DPN], b[DPN], un[DPN];
#pragma unroll
  for (int c = 0; c < DPN; ++c) {
    const int64_t r = r0 + c;
    const int64_t ci = coef_per_node ? v : r;  // alpha/beta per node when a node's DOFs share them
    a[c] = u1[r];
    b[c] = stream_load<NT>(u2 + r);
    const T ka = stream_load<NT>(alpha + ci) * acc[c];
    if constexpr (DAMPED) {
      const T be = pick_bits(T(2), beta[ci * (1 - undamped)], undamped);
      un[c] = be * a[c] + (T(1) - be) * b[c] - ka;
    } else {
      un[c] = (T(2) * a[c] - b[c]) - ka;
    }
  }
  for (int32_t i = L.fptr[lb], ie = L.fptr[lb + 1]; i < ie; ++i) {
    const int64_t rr = L.f_row[i];
    if (rr / DPN == v) {
      const T f = f_alpha[i] * F[static_cast<int64_t>(t) * F_rows + L.f_frow[i]];
#pragma unroll
      for (int c = 0; c < DPN; ++c)
        if (rr == l0 + c) un[c] += f;
    }
  }
  for (int32_t i = L.iptr[lb], ie = L.iptr[lb + 1]; i < ie; ++i) {
    const int64_t rr = L.i_row[i];
    if (rr / DPN == v) {
      const int32_t k = L.i_k[i];
      T c1 = T(0), c2 = T(0);
      for (int64_t j = I.cptr[k], je = I.cptr[k + 1]; j < je; ++j) {
        c1 += I.cval[j] * u1[I.ccol[j]];
        c2 += I.cval[j] * u2[I.ccol[j]];
      }
#pragma unroll
      for (int c = 0; c < DPN; ++c)
        if (rr == l0 + c) {
          un[c] += I.gfix[k] * b[c] + I.sq * c1 + I.sr * c2;
          if (I.wI != nullptr) I.wI[k] = un[c];
        }
    }
  }
#pragma unroll
  for (int c = 0; c < DPN; ++c) {
    stream_store<NT>(un[c], u + r0 + c);
    if (fld_prev != nullptr) fld_prev[r0 + c] = fld_is_vel ? (a[c] - b[c]) / dt : a[c];
  }
  if (hist_prev != nullptr)
    for (int32_t i = L.mptr[lb], ie = L.mptr[lb + 1]; i < ie; ++i) {
      const int64_t rr = L.m_row[i];
      if (rr / DPN == v) {
#pragma unroll
        for (int c = 0; c < DPN; ++c)
          if (rr == l0 + c) hist_prev[L.m_k[i]] = a[c];
      }
    }
}

// Slots in groups of G (n_slot is a multiple of G): a group's table entries,
// then its element records, are requested before any use, so a wave keeps G
// independent gathers in flight at each level of the table -> element -> u1
// chain.
template <typename T, int NPE, int DPN, int G>
__device__ __forceinline__ void ebe_gather(int64_t n_nodes, int64_t v, int32_t n_slot,
                                           const int32_t* __restrict__ adj,
                                           const int4* __restrict__ el_nodes,
                                           const T* __restrict__ el_scale,
                                           const T* __restrict__ srows, const T* __restrict__ u1,
                                           T* acc) {
  for (int32_t k0 = 0; k0 < n_slot; k0 += G) {
    int32_t e0[G];
    int4 nd[G];
    T s[G];
    ebe_load_table<T, G>(n_nodes, v, k0, adj, e0);
    ebe_load_records<T, G>(e0, el_nodes, el_scale, nd, s);
    ebe_group_apply<T, NPE, DPN, G>(k0, e0, nd, s, srows, u1, acc);
  }
}

template <typename T, bool DAMPED, int NPE, int DPN, int G>
__global__ __launch_bounds__(kBlock) void k_step_ebe(
    int64_t n_nodes, int64_t nblk, const int32_t* __restrict__ dof0, int32_t n_slot,
    const int32_t* __restrict__ adj, const int4* __restrict__ el_nodes,
    const T* __restrict__ el_scale,
    const T* __restrict__ srows, const T* __restrict__ alpha, const T* __restrict__ beta,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u, BlockLists L,
    const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows, int32_t t, int32_t q,
    T* __restrict__ hist_prev, T* __restrict__ fld_prev, int32_t fld_is_vel, T dt,
    IfaceRows<T> I, int32_t coef_per_node) {
  const int64_t lb = xcd_block(blockIdx.x, nblk);
  const int64_t v = lb * kBlock + threadIdx.x;
  if (v >= n_nodes) return;
  T acc[DPN];
#pragma unroll
  for (int c = 0; c < DPN; ++c) acc[c] = T(0);
  ebe_gather<T, NPE, DPN, G>(n_nodes, v, n_slot, adj, el_nodes, el_scale, srows, u1, acc);
  const int64_t r0 = dof0 != nullptr ? static_cast<int64_t>(dof0[v]) : v * DPN;
  ebe_finish<T, DAMPED, DPN>(v, lb, r0, acc, alpha, beta, u1, u2, u, L, f_alpha, F, F_rows, t,
                             q, hist_prev, fld_prev, fld_is_vel, dt, I, coef_per_node);
}

// ---------------------------------------------------------------------------
// Mesh patterns: the (slot, vertex) -> neighbour-class tables of the layout
// (slots and vertices in canonical order) of the structured meshes, so that
// k_step_ebe_pat can keep each neighbour's u1 in registers. Class SELF is the
// node itself.
struct PatKuhn24 {  // cubic cells split into 6 tetrahedra around a main diagonal
  static constexpr int ID = 1, NPE = 4, NS = 24, NC = 15, SELF = 7;
  static constexpr int8_t cls[NS][NPE] = {
      {0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7},
      {1, 3, 7, 8}, {1, 5, 7, 8}, {2, 3, 7, 9}, {2, 6, 7, 9}, {3, 7, 8, 10}, {3, 7, 9, 10},
      {4, 5, 7, 11}, {4, 6, 7, 11}, {5, 7, 8, 12}, {5, 7, 11, 12}, {6, 7, 9, 13},
      {6, 7, 11, 13}, {7, 8, 10, 14}, {7, 8, 12, 14}, {7, 9, 10, 14}, {7, 9, 13, 14},
      {7, 11, 12, 14}, {7, 11, 13, 14}};
  // Entries of slot k's template rows (row c, bit b*3 + d: canonical vertex
  // b, component d) that are not structurally zero: on axis-aligned cells
  // with an isotropic material, 420 of the 864 products of a node vanish
  // (the shape-function gradients of these tetrahedra have zero
  // components). A group takes the product over these entries only when
  // every other entry of its slot rows is zero or the rounding residue of
  // a cancellation in B^T D B (at most 1e-14 of the largest entry; about
  // 1e-17 on the cube), checked at set-up: the sums then agree with the
  // full product to rounding.
  static constexpr uint16_t nz[NS][3] = {
      {0x264, 0x5a0, 0x999}, {0x252, 0x5a9, 0x990}, {0x360, 0x4a4, 0x95a},
      {0x372, 0x489, 0x948}, {0x2d0, 0x4ec, 0x912}, {0x2f4, 0x4c8, 0x909},
      {0xb7a, 0x4bd, 0xb7a}, {0x6fc, 0x6fc, 0x93b}, {0x27e, 0xdb9, 0xdb9},
      {0x6fc, 0x6fc, 0x93b}, {0x5ed, 0xbd2, 0x5ed}, {0xdc9, 0x3f6, 0x3f6},
      {0x27e, 0xdb9, 0xdb9}, {0xb7a, 0x4bd, 0xb7a}, {0x9db, 0x9db, 0x7e4},
      {0xdc9, 0x3f6, 0x3f6}, {0x9db, 0x9db, 0x7e4}, {0x5ed, 0xbd2, 0x5ed},
      {0x999, 0x05a, 0x264}, {0x5a9, 0x252, 0x06c}, {0x099, 0x95a, 0x4a4},
      {0x489, 0x372, 0x0b4}, {0x129, 0x912, 0x4ec}, {0x909, 0x132, 0x2f4}};
  static constexpr bool nzbit(int k, int c, int col) { return (nz[k][c] >> col) & 1u; }
  // nxt[j]: the class whose node offset is class j's plus one (the next
  // node along the fastest mesh index), -1 if none; checked against the
  // group's offsets before a launch relies on it (k_step_ebe_multi)
  static constexpr int8_t nxt[NC] = {1, -1, 3, -1, 5, -1, 7, 8, -1, 10, -1, 12, -1, 14, -1};
};
struct PatTriRect {  // squares split into two right triangles
  static constexpr int ID = 2, NPE = 3, NS = 6, NC = 7, SELF = 3;
  static constexpr int8_t cls[NS][NPE] = {{0, 1, 3}, {0, 2, 3}, {1, 3, 4},
                                          {2, 3, 5}, {3, 4, 6}, {3, 5, 6}};
  static constexpr bool nzbit(int, int, int) { return true; }
  static constexpr int8_t nxt[NC] = {1, -1, 3, 4, -1, 6, -1};
};
struct PatTriIso {  // near-equilateral triangles
  static constexpr int ID = 3, NPE = 3, NS = 6, NC = 7, SELF = 2;
  static constexpr int8_t cls[NS][NPE] = {{0, 1, 2}, {0, 2, 3}, {1, 2, 4},
                                          {2, 3, 6}, {2, 4, 5}, {2, 5, 6}};
  static constexpr bool nzbit(int, int, int) { return true; }
  static constexpr int8_t nxt[NC] = {-1, -1, -1, -1, -1, -1, -1};  // no sliding window
};

// Whether a group's class offsets (in DOFs) follow a pattern's nxt table:
// off[nxt[j]] = off[j] + dpn wherever nxt[j] >= 0.
template <typename PT>
bool offsets_follow_next(const std::vector<int64_t>& doff, int dpn) {
  if (static_cast<int>(doff.size()) != PT::NC) return false;
  bool any = false;
  for (int j = 0; j < PT::NC; ++j)
    if (PT::nxt[j] >= 0) {
      if (doff[PT::nxt[j]] != doff[j] + dpn) return false;
      any = true;
    }
  return any;
}

// Class offsets (w - v) of a pattern's neighbours on regular nodes.
template <int NC>
struct ClassOffsets {
  int32_t off[NC];
};

// Where slot k's element keeps its scale in the anchor table (n_rows x
// n_nodes, row = the element's template, column = its canonical first
// vertex): entry v + base[k] on a regular node v, base[k] = row * n_nodes +
// the node offset of the first vertex; len = the table's length, below
// 2^31 (32-bit offsets: 64-bit ones cost the triangle kernels more scalar
// spills, 76 against 98 in the two-DOF multi-source kernel). Only
// patterns of few slots take it: the per-slot kernel arguments compete with
// the template rows for scalar registers (24 slots spill).
constexpr int kAnchorMaxSlots = 8;
template <int NS>
struct AnchorSlots {
  int32_t base[NS];
  int32_t len;
};

// Slots whose scale reads a batch of the slot product issues together (on
// the 24-slot tet pattern with the structurally sparse product, 24.3M DOF
// cube: 477 us per step with one slot at a time, 450 us with batches of 6 or
// 8, 445 us with 12, 927 us with one batch of 24, which spills; the dense
// product takes 8).
template <bool SP>
constexpr int slot_batch() { return SP ? 12 : 8; }

// The slot product of node v: acc += sum over occupied slots k of s_k R_k
// x_k, x_k the u1 values of slot k's vertices (neighbour classes cls[k][.])
// and s_k the element scale (1 unless bit k of sm). ib = 1 on an irregular
// node (its scaled slots read the slot -> element chain even in ANCHOR
// groups). Predicates are integer lane values (0 or 1) and bit masks: kept
// per lane in VGPRs, where 24 slots' worth of boolean lane masks would
// exhaust the scalar registers.
template <typename T, typename PT, int DPN, bool ANCHOR, bool REG, bool SP>
__device__ __forceinline__ void ebe_slot_sum(
    int64_t v, int64_t n_nodes, int64_t ib, uint32_t pm, uint32_t sm,
    const AnchorSlots<PT::NS>& an, const int32_t* __restrict__ adj,
    const T* __restrict__ el_scale, const T* __restrict__ anchor, const T* __restrict__ srows,
    const T (&x)[PT::NC][DPN], T (&acc)[DPN]) {
  constexpr int NPE = PT::NPE, NS = PT::NS, KE = NPE * DPN, kSlotBatch = slot_batch<SP>();
  if constexpr (!ANCHOR && NS % kSlotBatch == 0) {
    // slot -> element -> scale chains in batches: every table read of a
    // batch is issued, then every scale read, so a wave waits twice per
    // batch instead of twice per slot
#pragma unroll
    for (int k0 = 0; k0 < NS; k0 += kSlotBatch) {
      int32_t e[kSlotBatch];
#pragma unroll
      for (int i = 0; i < kSlotBatch; ++i) {
        const uint32_t sb = (sm >> (k0 + i)) & 1u;
        e[i] = adj[(static_cast<int64_t>(k0 + i) * n_nodes + v) * sb];
      }
      T s[kSlotBatch];
#pragma unroll
      for (int i = 0; i < kSlotBatch; ++i) {
        const uint32_t sb = (sm >> (k0 + i)) & 1u;
        s[i] = pick_bits(el_scale[static_cast<int64_t>(e[i]) * sb], T(1), sb);
      }
#pragma unroll
      for (int i = 0; i < kSlotBatch; ++i) {
        const int k = k0 + i;
        const T* R = srows + k * (DPN * KE);
        T y[DPN];
#pragma unroll
        for (int c = 0; c < DPN; ++c) y[c] = T(0);
#pragma unroll
        for (int b = 0; b < NPE; ++b)
#pragma unroll
          for (int d = 0; d < DPN; ++d)
#pragma unroll
            for (int c = 0; c < DPN; ++c)
              if (!SP || PT::nzbit(k, c, b * DPN + d))  // folds at compile time
                y[c] += R[c * KE + b * DPN + d] * x[PT::cls[k][b]][d];
        const uint32_t ob = (pm >> k) & 1u;
#pragma unroll
        for (int c = 0; c < DPN; ++c) acc[c] += pick_bits(s[i] * y[c], T(0), ob);
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const uint32_t sb = (sm >> k) & 1u;
    T s;
    if constexpr (ANCHOR) {
      // occupied slot on a regular node: its own entry; else any entry in
      // range (the slot's product is masked out)
      int64_t a = v + an.base[k];
      a = (a < 0 || a >= an.len) ? 0 : a;
      const uint32_t rb = sb & static_cast<uint32_t>(1 - ib);
      T sa = anchor[a * rb];
      if constexpr (!REG) {
        const uint32_t xb = sb & static_cast<uint32_t>(ib);
        const int32_t e = adj[(static_cast<int64_t>(k) * n_nodes + v) * xb];
        sa = pick_bits(el_scale[static_cast<int64_t>(e) * xb], sa, xb);
      }
      s = pick_bits(sa, T(1), sb);
    } else {
      const int32_t e = adj[(static_cast<int64_t>(k) * n_nodes + v) * sb];
      s = pick_bits(el_scale[static_cast<int64_t>(e) * sb], T(1), sb);
    }
    const T* R = srows + k * (DPN * KE);
    T y[DPN];
#pragma unroll
    for (int c = 0; c < DPN; ++c) y[c] = T(0);
#pragma unroll
    for (int b = 0; b < NPE; ++b)
#pragma unroll
      for (int d = 0; d < DPN; ++d)
#pragma unroll
        for (int c = 0; c < DPN; ++c)
          if (!SP || PT::nzbit(k, c, b * DPN + d))  // folds at compile time
            y[c] += R[c * KE + b * DPN + d] * x[PT::cls[k][b]][d];
    const uint32_t ob = (pm >> k) & 1u;
#pragma unroll
    for (int c = 0; c < DPN; ++c) acc[c] += pick_bits(s * y[c], T(0), ob);
  }
}

// Element-by-element step on a mesh pattern (at most 128 VGPRs, so that four
// wavefronts share a SIMD: the kernel is bound by load latency more than by
// its FP64 FMAs). Node v's first DOF is r0 (dof0[v], or v*DPN when dof0 is
// NULL). The NC neighbours' u1 values are gathered once and kept in
// registers; on a regular node the first DOF of neighbour j is r0 + off[j].
// Irregular nodes (bit 31 of pmask) read it from the class-major neighbour
// table nbr; in ANCHOR groups a wave that holds one takes the general path
// (REG = false), the others a path without those reads. pmask bit k marks an occupied slot,
// smask bit k a slot whose element scale differs from 1: only those read a
// scale, through slot table -> element -> scale. ANCHOR groups (most slots
// scaled, e.g. every triangle, whose scale carries J_rep / J_e) read a scale
// for every occupied slot, on a regular node from the anchor table, where
// each element's scale sits at (its template, its canonical first vertex): a
// wave's reads are coalesced and the element's other vertices find the line
// in cache. ANCHOR groups also read pmask, u2 and alpha and write u as
// non-temporal streams (on the 2D coupled model 103 -> 92 us per step; on
// the 3D tet mesh the same was 3 % slower). Template rows (canonical vertex
// order) come through the scalar cache.
template <typename T, bool DAMPED, typename PT, int DPN, bool ANCHOR, bool REG, bool SP = false>
__device__ __forceinline__ void ebe_pat_node(
    int64_t v, int64_t lb, int64_t n_nodes, int64_t n, int64_t r0, uint32_t pm, uint32_t sm,
    const ClassOffsets<PT::NC>& co, const AnchorSlots<PT::NS>& an,
    const int32_t* __restrict__ nbr, const int32_t* __restrict__ adj,
    const T* __restrict__ el_scale, const T* __restrict__ anchor, const T* __restrict__ srows,
    const T* __restrict__ stencil, const T* __restrict__ alpha, const T* __restrict__ beta,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u, const BlockLists& L,
    const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows, int32_t t, int32_t q,
    T* __restrict__ hist_prev, T* __restrict__ fld_prev, int32_t fld_is_vel, T dt,
    const IfaceRows<T>& I, int32_t coef_per_node) {
  constexpr int NS = PT::NS, NC = PT::NC;
  constexpr uint32_t kAll = (1u << NS) - 1u;
  const int64_t ib = REG ? 0 : static_cast<int64_t>(pm >> 31);  // 1: irregular node
  // first DOF of neighbour j (an offset out of range belongs to an absent
  // neighbour, used only by empty slots)
  auto nb = [&](int j) -> int64_t {
    if (j == PT::SELF) return r0;
    const int64_t wr = r0 + co.off[j];
    const int64_t wreg = (wr < 0 || wr > n - DPN) ? r0 : wr;
    if constexpr (REG) {
      return wreg;
    } else {
      // lanes that do not need a table entry read entry 0 instead, one
      // cached line for the wave (no divergent branch around the load)
      const int32_t wj = nbr[(static_cast<int64_t>(j) * n_nodes + v) * ib];
      const int64_t wi = wj < 0 ? r0 : wj;
      return ib * wi + (1 - ib) * wreg;
    }
  };
  T acc[DPN];
#pragma unroll
  for (int c = 0; c < DPN; ++c) acc[c] = T(0);
  if (!ANCHOR && __all((pm & kAll) == kAll && sm == 0u)) {
    // every lane's elements are all present with scale 1 (the interior): K u
    // is the fixed stencil sum_j B_j u_j, B_j the sum of the template blocks
    // of neighbour class j (NC DPN x DPN blocks instead of NS template rows)
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const int64_t w = nb(j);
#pragma unroll
      for (int d = 0; d < DPN; ++d) {
        const T xd = u1[w + d];
#pragma unroll
        for (int c = 0; c < DPN; ++c) acc[c] += stencil[(j * DPN + c) * DPN + d] * xd;
      }
    }
    ebe_finish<T, DAMPED, DPN, ANCHOR>(v, lb, r0, acc, alpha, beta, u1, u2, u, L, f_alpha, F,
                                       F_rows, t, q, hist_prev, fld_prev, fld_is_vel, dt, I,
                                       coef_per_node, (pm >> 30) & 1u);
    return;
  }
  T x[NC][DPN];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const int64_t w = nb(j);
#pragma unroll
    for (int d = 0; d < DPN; ++d) x[j][d] = u1[w + d];
  }
  ebe_slot_sum<T, PT, DPN, ANCHOR, REG, SP>(v, n_nodes, ib, pm, sm, an, adj, el_scale, anchor, srows,
                                             x, acc);
  ebe_finish<T, DAMPED, DPN, ANCHOR>(v, lb, r0, acc, alpha, beta, u1, u2, u, L, f_alpha, F,
                                     F_rows, t, q, hist_prev, fld_prev, fld_is_vel, dt, I,
                                     coef_per_node, (pm >> 30) & 1u);
}

template <typename T, bool DAMPED, typename PT, int DPN, bool ANCHOR, bool SP = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_step_ebe_pat(
    int64_t n_nodes, int64_t nblk, int64_t n, const int32_t* __restrict__ dof0,
    ClassOffsets<PT::NC> co, AnchorSlots<PT::NS> an, const uint32_t* __restrict__ pmask,
    const uint32_t* __restrict__ smask, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ adj, const T* __restrict__ el_scale,
    const T* __restrict__ anchor,
    const T* __restrict__ srows, const T* __restrict__ stencil, const T* __restrict__ alpha,
    const T* __restrict__ beta, const T* __restrict__ u1, const T* __restrict__ u2,
    T* __restrict__ u, BlockLists L, const T* __restrict__ f_alpha, const T* __restrict__ F,
    int32_t F_rows, int32_t t, int32_t q, T* __restrict__ hist_prev, T* __restrict__ fld_prev,
    int32_t fld_is_vel, T dt, IfaceRows<T> I, int32_t coef_per_node) {
  constexpr uint32_t kAll = (1u << PT::NS) - 1u;
  const int64_t lb = xcd_block(blockIdx.x, nblk);
  const int64_t v = lb * kBlock + threadIdx.x;
  if (v >= n_nodes) return;
  const uint32_t pm = stream_load<ANCHOR>(pmask + v);
  const uint32_t sm = ANCHOR ? (pm & kAll) : smask[v];
  const int64_t r0 = dof0 != nullptr ? static_cast<int64_t>(dof0[v]) : v * DPN;
  // (the split pays on the anchored triangle groups; on the 24-slot tet
  // pattern the second copy costs more scalar registers than it saves)
  if (!ANCHOR || __any(static_cast<int>(pm >> 31)))
    ebe_pat_node<T, DAMPED, PT, DPN, ANCHOR, false, SP>(
        v, lb, n_nodes, n, r0, pm, sm, co, an, nbr, adj, el_scale, anchor, srows, stencil, alpha,
        beta, u1, u2, u, L, f_alpha, F, F_rows, t, q, hist_prev, fld_prev, fld_is_vel, dt, I,
        coef_per_node);
  else
    ebe_pat_node<T, DAMPED, PT, DPN, ANCHOR, true>(
        v, lb, n_nodes, n, r0, pm, sm, co, an, nbr, adj, el_scale, anchor, srows, stencil, alpha,
        beta, u1, u2, u, L, f_alpha, F, F_rows, t, q, hist_prev, fld_prev, fld_is_vel, dt, I,
        coef_per_node);
}

// ---------------------------------------------------------------------------
// Multi-source element-by-element step (the element product of the batched
// march): u is n x S (source axis contiguous); one wavefront per node, lane =
// source, so everything about the node (its masks, first DOF, element
// scales, template rows) is wave-uniform and read through the scalar cache,
// and each neighbour row of u1 is one coalesced S-value gather. A wavefront
// marches npw consecutive nodes; along a run of regular nodes whose first
// DOFs advance by DPN, the NC neighbour rows held in registers slide by one
// node (x[j] <- x[nxt[j]]) and only the classes without a successor are
// gathered: 3 of 7 node rows for triangles, 7 of 15 for the tets. The update,
// forcing (f_src = lane), interface and monitor handling are those of
// k_step_multi on the node's rows; the lists hold the group's local rows
// v*DPN + c and are grouped per workgroup of this launch.

// K u1 on node v's DPN rows for every lane (source): the slot product, or
// the fixed stencil where every element is present with scale 1. X(j, d) is
// the u1 value of neighbour class j, component d.
template <typename T, typename PT, int DPN, bool ANCHOR, bool SP, typename XF>
__device__ __forceinline__ void ebe_multi_sum(
    int64_t v, int64_t n_nodes, uint32_t pm, uint32_t sm, const AnchorSlots<PT::NS>& an,
    const int32_t* __restrict__ adj, const T* __restrict__ el_scale,
    const T* __restrict__ anchor, const T* __restrict__ srows, const T* __restrict__ stencil,
    XF X, T (&acc)[DPN]) {
  constexpr int NPE = PT::NPE, NS = PT::NS, NC = PT::NC, KE = NPE * DPN;
  constexpr uint32_t kAll = (1u << NS) - 1u;
#pragma unroll
  for (int c = 0; c < DPN; ++c) acc[c] = T(0);
  if (!ANCHOR && (pm & kAll) == kAll && sm == 0u) {
    // every element present with scale 1: the fixed stencil sum_j B_j u_j
#pragma unroll
    for (int j = 0; j < NC; ++j)
#pragma unroll
      for (int d = 0; d < DPN; ++d)
#pragma unroll
        for (int c = 0; c < DPN; ++c) acc[c] += stencil[(j * DPN + c) * DPN + d] * X(j, d);
    return;
  }
  const bool irr = (pm >> 31) != 0u;
  // every slot's scale read is issued before any product (wave-uniform
  // scalar loads, branch-free: a slot that needs none reads entry 0 and
  // keeps 1), so the node waits for them once instead of once per slot
  T sc[NS];
  if (ANCHOR && !irr) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int64_t a = v + an.base[k];
      const bool use = ((sm >> k) & 1u) && a >= 0 && a < an.len;
      const T x = anchor[use ? a : 0];
      sc[k] = use ? x : T(1);
    }
  } else {
    int32_t el[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k)
      el[k] = ((sm >> k) & 1u) ? adj[static_cast<int64_t>(k) * n_nodes + v] : 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const T x = el_scale[el[k]];
      sc[k] = ((sm >> k) & 1u) ? x : T(1);
    }
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    if (((pm >> k) & 1u) == 0u) continue;  // empty slot (uniform branch)
    const T s = sc[k];
    const T* R = srows + k * (DPN * KE);
    T y[DPN];
#pragma unroll
    for (int c = 0; c < DPN; ++c) y[c] = T(0);
#pragma unroll
    for (int b = 0; b < NPE; ++b)
#pragma unroll
      for (int d = 0; d < DPN; ++d)
#pragma unroll
        for (int c = 0; c < DPN; ++c)
          if (!SP || PT::nzbit(k, c, b * DPN + d))
            y[c] += R[c * KE + b * DPN + d] * X(PT::cls[k][b], d);
#pragma unroll
    for (int c = 0; c < DPN; ++c) acc[c] += s * y[c];
  }
}

// The update of node v's rows r0 .. r0 + DPN - 1 for one lane (k_step_multi's
// per row): a = u1 and b = u2 of the rows, acc = (K u1) of the rows.
template <typename T, bool DAMPED, int DPN>
__device__ __forceinline__ void ebe_multi_update(
    int64_t lb, int64_t v, int64_t r0, int lane, int32_t S, uint32_t pm, const T (&acc)[DPN],
    const T (&a)[DPN], const T (&b)[DPN], const T* __restrict__ alpha,
    const T* __restrict__ beta, int32_t coef_per_node, const T* __restrict__ u1,
    const T* __restrict__ u2, T* __restrict__ u, const BlockLists& L,
    const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows, int32_t t,
    int32_t q, T* __restrict__ hist_prev, const IfaceRows<T>& I) {
  const uint32_t undamped = (pm >> 30) & 1u;
  const int64_t l0 = v * DPN;
  T un[DPN];
#pragma unroll
  for (int c = 0; c < DPN; ++c) {
    const int64_t ci = coef_per_node ? v : r0 + c;
    const T ka = alpha[ci] * acc[c];
    if constexpr (DAMPED) {
      const T be = undamped ? T(2) : beta[ci];
      un[c] = be * a[c] + (T(1) - be) * b[c] - ka;
    } else {
      un[c] = (T(2) * a[c] - b[c]) - ka;
    }
  }
  for (int32_t i = L.fptr[lb], ie = L.fptr[lb + 1]; i < ie; ++i) {
    const int64_t rr = L.f_row[i];
    if (rr / DPN == v && L.f_src[i] == lane) {
      const T f = f_alpha[i] * F[static_cast<int64_t>(t) * F_rows + L.f_frow[i]];
#pragma unroll
      for (int c = 0; c < DPN; ++c)
        if (rr == l0 + c) un[c] += f;
    }
  }
  for (int32_t i = L.iptr[lb], ie = L.iptr[lb + 1]; i < ie; ++i) {
    const int64_t rr = L.i_row[i];
    if (rr / DPN != v) continue;
    const int32_t k = L.i_k[i];
    T c1 = T(0), c2 = T(0);
    for (int64_t p = I.cptr[k], pe = I.cptr[k + 1]; p < pe; ++p) {
      const int64_t cc = static_cast<int64_t>(I.ccol[p]) * S + lane;
      c1 += I.cval[p] * u1[cc];
      c2 += I.cval[p] * u2[cc];
    }
#pragma unroll
    for (int c = 0; c < DPN; ++c)
      if (rr == l0 + c) {
        un[c] += I.gfix[k] * b[c] + I.sq * c1 + I.sr * c2;
        if (I.wI != nullptr) I.wI[static_cast<int64_t>(k) * S + lane] = un[c];
      }
  }
#pragma unroll
  for (int c = 0; c < DPN; ++c) stream_store<true>(un[c], u + (r0 + c) * S + lane);
  if (hist_prev != nullptr)
    for (int32_t i = L.mptr[lb], ie = L.mptr[lb + 1]; i < ie; ++i) {
      const int64_t rr = L.m_row[i];
      if (rr / DPN == v) {
#pragma unroll
        for (int c = 0; c < DPN; ++c)
          if (rr == l0 + c) hist_prev[L.m_k[i] + static_cast<int64_t>(q) * lane] = a[c];
      }
    }
}

// K u1 on node v's DPN rows from the node's neighbour-class blocks
// (node-block form): nbv holds the NC blocks of DPN x DPN, class-major.
template <typename T, typename PT, int DPN, typename XF>
__device__ __forceinline__ void node_block_sum(const T* __restrict__ nbv, XF X, T (&acc)[DPN]) {
#pragma unroll
  for (int c = 0; c < DPN; ++c) acc[c] = T(0);
#pragma unroll
  for (int j = 0; j < PT::NC; ++j)
#pragma unroll
    for (int c = 0; c < DPN; ++c)
#pragma unroll
      for (int d = 0; d < DPN; ++d) acc[c] += nbv[(j * DPN + c) * DPN + d] * X(j, d);
}

// Nodes vbeg .. vend - 1 of one wavefront, one after the other (the sliding
// neighbour rows along runs of regular nodes when slide is set). NB: K u1
// from the node-block form nb (n_nodes x NC x DPN x DPN) instead of the slot
// product.
template <typename T, bool DAMPED, typename PT, int DPN, bool ANCHOR, bool SP, bool NB = false>
__device__ __forceinline__ void ebe_multi_nodes(
    int64_t lb, int64_t vbeg, int64_t vend, int lane, int32_t S, int32_t slide, int64_t n_nodes,
    int64_t n, const int32_t* __restrict__ dof0, const ClassOffsets<PT::NC>& co,
    const AnchorSlots<PT::NS>& an, const uint32_t* __restrict__ pmask,
    const uint32_t* __restrict__ smask, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ adj, const T* __restrict__ el_scale,
    const T* __restrict__ anchor, const T* __restrict__ srows, const T* __restrict__ stencil,
    const T* __restrict__ alpha, const T* __restrict__ beta, const T* __restrict__ u1,
    const T* __restrict__ u2, T* __restrict__ u, const BlockLists& L,
    const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows, int32_t t,
    int32_t q, T* __restrict__ hist_prev, const IfaceRows<T>& I, int32_t coef_per_node,
    const T* __restrict__ nb) {
  constexpr int NS = PT::NS, NC = PT::NC;
  constexpr uint32_t kAll = (1u << NS) - 1u;
  const bool active = lane < S;
  auto load = [&](int64_t r) -> T { return active ? u1[r * S + lane] : T(0); };
  // first DOF of neighbour j on a regular node (an offset out of range
  // belongs to an absent neighbour, used only by empty slots)
  auto reg = [&](int j, int64_t r0) -> int64_t {
    const int64_t wr = r0 + co.off[j];
    return (wr < 0 || wr > n - DPN) ? r0 : wr;
  };
  T x[NC][DPN];
  int64_t rprev = -1;
  bool win = false;
  for (int64_t v = vbeg; v < vend; ++v) {
    const uint32_t pm = pmask[v];
    const uint32_t sm = ANCHOR ? (pm & kAll) : smask[v];
    const int64_t r0 = dof0 != nullptr ? static_cast<int64_t>(dof0[v]) : v * DPN;
    // u2 of the node's rows, issued with the neighbour gathers (one wait per node)
    T b[DPN];
#pragma unroll
    for (int c = 0; c < DPN; ++c) b[c] = active ? stream_load<true>(u2 + (r0 + c) * S + lane) : T(0);
    if (pm >> 31) {  // irregular node: neighbours from the class-major table
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        int64_t w = r0;
        if (j != PT::SELF) {
          const int32_t wj = nbr[static_cast<int64_t>(j) * n_nodes + v];
          w = wj < 0 ? r0 : wj;
        }
#pragma unroll
        for (int d = 0; d < DPN; ++d) x[j][d] = load(w + d);
      }
      win = false;
    } else if (win && slide && r0 == rprev + DPN) {
#pragma unroll
      for (int j = 0; j < NC; ++j) {  // ascending: nxt[j] > j still holds the old row
        // (j is a constant once unrolled: the branch and the index fold)
        constexpr int8_t const* nx = PT::nxt;
        if (nx[j] >= 0) {
#pragma unroll
          for (int d = 0; d < DPN; ++d) x[j][d] = x[nx[j] >= 0 ? nx[j] : j][d];
        } else {
          const int64_t w = reg(j, r0);
#pragma unroll
          for (int d = 0; d < DPN; ++d) x[j][d] = load(w + d);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const int64_t w = j == PT::SELF ? r0 : reg(j, r0);
#pragma unroll
        for (int d = 0; d < DPN; ++d) x[j][d] = load(w + d);
      }
      win = true;
    }
    rprev = r0;
    T acc[DPN];
    if constexpr (NB)
      node_block_sum<T, PT, DPN>(nb + v * (NC * DPN * DPN), [&](int j, int d) { return x[j][d]; },
                                 acc);
    else
      ebe_multi_sum<T, PT, DPN, ANCHOR, SP>(v, n_nodes, pm, sm, an, adj, el_scale, anchor,
                                            srows, stencil,
                                            [&](int j, int d) { return x[j][d]; }, acc);
    if (!active) continue;
    T a[DPN];
#pragma unroll
    for (int c = 0; c < DPN; ++c) a[c] = x[PT::SELF][c];
    ebe_multi_update<T, DAMPED, DPN>(lb, v, r0, lane, S, pm, acc, a, b, alpha, beta,
                                     coef_per_node, u1, u2, u, L, f_alpha, F, F_rows, t, q,
                                     hist_prev, I);
  }
}

template <typename T, bool DAMPED, typename PT, int DPN, bool ANCHOR, bool SP = false,
          bool NB = false>
__global__ __launch_bounds__(kBlock) void k_step_ebe_multi(
    int64_t n_nodes, int64_t nblk, int32_t npw, int32_t S, int32_t slide, int64_t n,
    const int32_t* __restrict__ dof0, ClassOffsets<PT::NC> co, AnchorSlots<PT::NS> an,
    const uint32_t* __restrict__ pmask, const uint32_t* __restrict__ smask,
    const int32_t* __restrict__ nbr, const int32_t* __restrict__ adj,
    const T* __restrict__ el_scale, const T* __restrict__ anchor, const T* __restrict__ srows,
    const T* __restrict__ stencil, const T* __restrict__ alpha, const T* __restrict__ beta,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u, BlockLists L,
    const T* __restrict__ f_alpha, const T* __restrict__ F, int32_t F_rows, int32_t t, int32_t q,
    T* __restrict__ hist_prev, IfaceRows<T> I, int32_t coef_per_node, const T* __restrict__ nb) {
  const int64_t lb = xcd_block(blockIdx.x, nblk);
  const int64_t wave =
      lb * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int64_t vbeg = wave * npw;
  if (vbeg >= n_nodes) return;  // uniform across the wavefront
  const int64_t vend = min(vbeg + npw, n_nodes);
  ebe_multi_nodes<T, DAMPED, PT, DPN, ANCHOR, SP, NB>(
      lb, vbeg, vend, lane, S, slide, n_nodes, n, dof0, co, an, pmask, smask, nbr, adj, el_scale,
      anchor, srows, stencil, alpha, beta, u1, u2, u, L, f_alpha, F, F_rows, t, q, hist_prev, I,
      coef_per_node, nb);
}

// Node-block form of one element group: block (j, c, d) of node v is the sum
// over v's occupied slots k and vertices b of class j of s_k R_k[c, b DPN + d]
// (the slot product of ebe_multi_sum with the u1 values left out), so
// K u1 on node v's rows is sum_j B_j(v) u1(class j). One thread per node.
template <typename T, typename PT, int DPN, bool ANCHOR>
__global__ __launch_bounds__(kBlock) void k_node_blocks(
    int64_t n_nodes, AnchorSlots<PT::NS> an, const uint32_t* __restrict__ pmask,
    const uint32_t* __restrict__ smask, const int32_t* __restrict__ adj,
    const T* __restrict__ el_scale, const T* __restrict__ anchor, const T* __restrict__ srows,
    T* __restrict__ nb) {
  constexpr int NPE = PT::NPE, NS = PT::NS, NC = PT::NC, KE = NPE * DPN, BS = NC * DPN * DPN;
  constexpr uint32_t kAll = (1u << NS) - 1u;
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n_nodes) return;
  const uint32_t pm = pmask[v];
  const uint32_t sm = ANCHOR ? (pm & kAll) : smask[v];
  const bool irr = (pm >> 31) != 0u;
  T B[BS];
#pragma unroll
  for (int i = 0; i < BS; ++i) B[i] = T(0);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    if (((pm >> k) & 1u) == 0u) continue;
    T s = T(1);
    if ((sm >> k) & 1u) {
      if (ANCHOR && !irr) {
        const int64_t a = v + an.base[k];
        s = (a < 0 || a >= an.len) ? T(1) : anchor[a];
      } else {
        s = el_scale[adj[static_cast<int64_t>(k) * n_nodes + v]];
      }
    }
    const T* R = srows + k * (DPN * KE);
#pragma unroll
    for (int b = 0; b < NPE; ++b) {
      const int j = PT::cls[k][b];
#pragma unroll
      for (int c = 0; c < DPN; ++c)
#pragma unroll
        for (int d = 0; d < DPN; ++d) B[(j * DPN + c) * DPN + d] += s * R[c * KE + b * DPN + d];
    }
  }
#pragma unroll
  for (int i = 0; i < BS; ++i) nb[v * BS + i] = B[i];
}

// Pattern id of a layout's class table, 0 if none matches.
int match_pattern(const bfe::EbeLayout& L) {
  auto same = [&](auto pat) {
    using P = decltype(pat);
    if (L.npe != P::NPE || L.n_slot != P::NS || L.n_cls != P::NC || L.self_cls != P::SELF)
      return false;
    for (int k = 0; k < P::NS; ++k)
      for (int b = 0; b < P::NPE; ++b)
        if (L.cls[k * P::NPE + b] != P::cls[k][b]) return false;
    return true;
  };
  if (L.n_cls == 0) return 0;
  if (same(PatKuhn24{})) return PatKuhn24::ID;
  if (same(PatTriRect{})) return PatTriRect::ID;
  if (same(PatTriIso{})) return PatTriIso::ID;
  return 0;
}

// ---------------------------------------------------------------------------
// Per-element kinetic-energy field of code/internal/fn_get_field_output.m:64-79
// from velocity frames (n x F): KE_node = sum over the node's DOFs of
// |v^2 m| / 2, then the mean over the element's nodes (absent: -1), then the
// square root. One thread per (element, frame).
template <typename T>
__global__ void k_field_ke(int64_t n_el, int32_t npe, const int32_t* __restrict__ els,
                           int32_t max_df, const int32_t* __restrict__ node_dof,
                           const T* __restrict__ m, const T* __restrict__ fld, int64_t n,
                           int32_t F, T* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_el * F) return;
  const int64_t e = i % n_el, f = i / n_el;
  const T* v = fld + f * n;
  T sum = T(0);
  int cnt = 0;
  for (int a = 0; a < npe; ++a) {
    const int32_t node = els[e * npe + a];
    if (node < 0) continue;
    ++cnt;
    T ke = T(0);
    for (int d = 0; d < max_df; ++d) {
      const int32_t g = node_dof[static_cast<int64_t>(node) * max_df + d];
      if (g >= 0) ke += fabs(v[g] * v[g] * m[g]) / T(2);
    }
    sum += ke;
  }
  out[i] = cnt > 0 ? sqrt(sum / T(cnt)) : T(0);
}

// The other per-element fields of code/internal/fn_get_field_output.m:28-58
// from displacement or velocity frames (n x F), one thread per (element,
// frame). BFE_FIELD_MEAN: the mean over the element's nodes of DOF column
// comp (an absent DOF counts as 0, :35-37 and :118-134). BFE_FIELD_DIV /
// _CURL: the sum over the element's nodes of the node's components 1..dim
// dotted with the unit vector from the element centre to the node (:85-98;
// for CURL turned to (y, -x), :100-105); 0 on an element with an absent
// node (:95). Centres, lengths and sums in the reference's order, without
// FMA contraction, so the host restatement's values come out bit for bit.
template <typename T>
__global__ void k_field_reduce(int32_t type, int32_t comp, int64_t n_el, int32_t npe,
                               const int32_t* __restrict__ els, int32_t max_df,
                               const int32_t* __restrict__ node_dof,
                               const double* __restrict__ nds, int32_t dim,
                               const T* __restrict__ fld, int64_t n, int32_t F,
                               T* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_el * F) return;
  const int64_t e = i % n_el, f = i / n_el;
  const T* v = fld + f * n;
  const int32_t* en = els + e * npe;
  auto val = [&](int32_t node, int d) -> T {
    const int32_t g = node_dof[static_cast<int64_t>(node) * max_df + d];
    return g >= 0 ? v[g] : T(0);
  };
  if (type == BFE_FIELD_MEAN) {
    T sum = T(0);
    int cnt = 0;
    for (int a = 0; a < npe; ++a) {
      if (en[a] < 0) continue;
      sum += val(en[a], comp);
      ++cnt;
    }
    out[i] = sum / T(cnt);
    return;
  }
  for (int a = 0; a < npe; ++a)
    if (en[a] < 0) {
      out[i] = T(0);
      return;
    }
  double c[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < npe; ++a)
    for (int d = 0; d < dim; ++d) c[d] += nds[static_cast<int64_t>(en[a]) * dim + d];
  for (int d = 0; d < dim; ++d) c[d] /= static_cast<double>(npe);
  T sum = T(0);
  for (int a = 0; a < npe; ++a) {
    double w[3], len = 0.0;
    for (int d = 0; d < dim; ++d) {
      w[d] = nds[static_cast<int64_t>(en[a]) * dim + d] - c[d];
      len += w[d] * w[d];
    }
    len = sqrt(len);
    for (int d = 0; d < dim; ++d) w[d] /= len;
    if (type == BFE_FIELD_CURL) {
      const double x = w[0];
      w[0] = w[1];
      w[1] = -x;
    }
    T dot = T(0);
    for (int d = 0; d < dim; ++d) dot += val(en[a], d) * static_cast<T>(w[d]);
    sum += dot;
  }
  out[i] = sum;
}

// ---------------------------------------------------------------------------
// Per-element gradients over one checkpoint segment of Ls time steps.
// Thread e keeps K_e (KMAX x KMAX), the lumped mass per unit density and the
// last two forward states of its DOFs in registers and walks the segment:
//   gK_e -= dt^2 * sum_t lam_t' K_e u_{t-1}
//   gM_e -= sum_t lam_t' Mbar_e (w0 u_t + w1 u_{t-1} + w2 u_{t-2})
// U and Lam hold the segment's forward and adjoint states (Ls x n, column
// t - segment start); um1/um2 are the two states before the segment.
template <typename T, int KMAX>
__global__ __launch_bounds__(kBlock) void k_grad(
    int32_t n_el, const int32_t* __restrict__ el_dofs, const T* __restrict__ el_K,
    const T* __restrict__ el_m, const T* __restrict__ el_w, int64_t n,
    const T* __restrict__ U, const T* __restrict__ Lam, int32_t Ls,
    const T* __restrict__ um1, const T* __restrict__ um2, double dt2,
    double* __restrict__ gK, double* __restrict__ gM) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_el) return;
  int32_t d[KMAX];
  T Ke[KMAX * KMAX], me[KMAX], p1[KMAX], p2[KMAX];
#pragma unroll
  for (int i = 0; i < KMAX; ++i) {
    d[i] = el_dofs[e * KMAX + i];
    me[i] = el_m[e * KMAX + i];
    const int64_t di = d[i] >= 0 ? d[i] : 0;  // unused slots read DOF 0, then discard it
    const T a1 = um1[di], a2 = um2[di];
    p1[i] = d[i] >= 0 ? a1 : T(0);
    p2[i] = d[i] >= 0 ? a2 : T(0);
  }
#pragma unroll
  for (int i = 0; i < KMAX * KMAX; ++i) Ke[i] = el_K[e * KMAX * KMAX + i];
  const T w0 = el_w[3 * e], w1 = el_w[3 * e + 1], w2 = el_w[3 * e + 2];
  double sk = 0.0, sm = 0.0;
  // two steps' gathers in flight (adjoint 1397 -> 1379 ms on 361k DOF x 16
  // transmitters; four took all 256 VGPRs and ran slower): the sums keep
  // their order
#pragma unroll 2
  for (int32_t t = 0; t < Ls; ++t) {
    T ut[KMAX], lt[KMAX];
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      const int64_t o = static_cast<int64_t>(t) * n + (d[i] >= 0 ? d[i] : 0);
      const T x = U[o], y = Lam[o];
      ut[i] = d[i] >= 0 ? x : T(0);
      lt[i] = d[i] >= 0 ? y : T(0);
    }
    T ks = T(0), ms = T(0);
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      T r = T(0);
#pragma unroll
      for (int j = 0; j < KMAX; ++j) r += Ke[i * KMAX + j] * p1[j];
      ks += lt[i] * r;
      ms += lt[i] * me[i] * (w0 * ut[i] + w1 * p1[i] + w2 * p2[i]);
    }
    sk += static_cast<double>(ks);
    sm += static_cast<double>(ms);
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      p2[i] = p1[i];
      p1[i] = ut[i];
    }
  }
  gK[e] -= dt2 * sk;
  gM[e] -= sm;
}

// k_grad over S sources at once: one thread per (element, source), GS
// (a power of two >= S) consecutive lanes per element, so the lanes of an
// element read the S sources of each DOF as one contiguous run. State
// (t, dof, s) of the forward march sits at U[t * us + dof * ud + s], of the
// adjoint at Lam[t * ls + dof * ld + s] (ls may be negative: the adjoint is
// stored in reverse time); the states before the segment are n x S. The
// element's sums over its sources are reduced across the GS lanes and added
// by the first lane. (Splitting the time loop into blocks of 16 steps, with
// consecutive workgroups on one block, ran 27.0 -> 28.2 ms per launch on the
// 16-transmitter adjoint workload: not kept.)
template <typename T, int KMAX>
__global__ __launch_bounds__(kBlock) void k_grad_multi(
    int32_t n_el, int32_t S, int32_t GS, const int32_t* __restrict__ el_dofs,
    const T* __restrict__ el_K, const T* __restrict__ el_m, const T* __restrict__ el_w,
    const T* __restrict__ U, int64_t us, int64_t ud, const T* __restrict__ Lam, int64_t ls,
    int64_t ld, int32_t Ls, const T* __restrict__ um1, const T* __restrict__ um2, double dt2,
    double* __restrict__ gK, double* __restrict__ gM) {
  const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t e = tid / GS;
  const int s = static_cast<int>(tid - e * GS);
  const bool on = e < n_el && s < S;
  const int64_t ee = e < n_el ? e : 0;  // threads past the end read element 0, then drop it
  const int ss = s < S ? s : 0;
  int32_t d[KMAX];
  T Ke[KMAX * KMAX], me[KMAX], p1[KMAX], p2[KMAX];
#pragma unroll
  for (int i = 0; i < KMAX; ++i) {
    d[i] = el_dofs[ee * KMAX + i];
    me[i] = el_m[ee * KMAX + i];
    const int64_t di = d[i] >= 0 ? d[i] : 0;
    const T a1 = um1[di * S + ss], a2 = um2[di * S + ss];
    p1[i] = d[i] >= 0 ? a1 : T(0);
    p2[i] = d[i] >= 0 ? a2 : T(0);
  }
#pragma unroll
  for (int i = 0; i < KMAX * KMAX; ++i) Ke[i] = el_K[ee * KMAX * KMAX + i];
  const T w0 = el_w[3 * ee], w1 = el_w[3 * ee + 1], w2 = el_w[3 * ee + 2];
  double sk = 0.0, sm = 0.0;
  for (int32_t t = 0; t < Ls; ++t) {
    T ut[KMAX], lt[KMAX];
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      const int64_t di = d[i] >= 0 ? d[i] : 0;
      const T x = U[t * us + di * ud + ss], y = Lam[t * ls + di * ld + ss];
      ut[i] = d[i] >= 0 ? x : T(0);
      lt[i] = d[i] >= 0 ? y : T(0);
    }
    T ks = T(0), ms = T(0);
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      T r = T(0);
#pragma unroll
      for (int j = 0; j < KMAX; ++j) r += Ke[i * KMAX + j] * p1[j];
      ks += lt[i] * r;
      ms += lt[i] * me[i] * (w0 * ut[i] + w1 * p1[i] + w2 * p2[i]);
    }
    sk += static_cast<double>(ks);
    sm += static_cast<double>(ms);
#pragma unroll
    for (int i = 0; i < KMAX; ++i) {
      p2[i] = p1[i];
      p1[i] = ut[i];
    }
  }
  if (!on) sk = sm = 0.0;
  for (int o = GS >> 1; o > 0; o >>= 1) {  // the element's lanes, in one wavefront
    sk += __shfl_xor(sk, o, 64);
    sm += __shfl_xor(sm, o, 64);
  }
  if (s == 0 && e < n_el) {
    gK[e] -= dt2 * sk;
    gM[e] -= sm;
  }
}

// ---------------------------------------------------------------------------
// Gradients of the batched adjoint as node-pair correlations. With element
// vertices grouped into nodes of GD consecutive DOFs, the stiffness term of
// element e is
//   sum_t lam_t' K_e u_{t-1} = sum_{i,j in e} K_e[i, j] C[d_i, d_j],
//   C[a, b] = sum_{t, s} lam_a(t, s) u_b(t - 1, s),
// and the mass term needs D0[a] = sum lam_a(t) u_a(t), C[a, a] and
// D2[a] = sum lam_a(t) u_a(t - 2). C is needed only on node pairs that share
// an element, so k_pair_acc gives one thread per (node a, source s) that
// walks the segment once, reading node a's and its EM upper neighbours' u and
// lam rows (16 rows per step for 2D solids against the 24 that the two
// elements per node read in k_grad_multi) and accumulating in FP64:
//   self block C[a, a] (GD x GD), D0[a], D2[a] and, for each upper
//   neighbour b (b > a, k-th in a's list), C[a, b] and C[b, a].
// Sums over the sources are reduced across the node's GS lanes and added to
// node a's record (pair_stride doubles) by its first lane; records persist
// over the segments, and k_pair_contract forms every element's gradient from
// them once after the sweep. Absent neighbours (slot k beyond a's list) read
// node a's own rows and their sums are never used.
template <int GD, int EM>
constexpr int pair_stride() {
  return GD * GD + 2 * GD + EM * 2 * GD * GD;
}

template <typename T, int GD, int EM>
__global__ __launch_bounds__(kBlock) void k_pair_acc(
    int64_t n_groups, int32_t S, int32_t GS, const int32_t* __restrict__ gdof,
    const int32_t* __restrict__ gnb, const T* __restrict__ U, int64_t us, int64_t ud,
    const T* __restrict__ Lam, int64_t ls, int64_t ld, int32_t Ls,
    const T* __restrict__ um1, const T* __restrict__ um2, double* __restrict__ acc) {
  constexpr int BS = GD * GD, STRIDE = pair_stride<GD, EM>();
  const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t a = tid / GS;
  const int s = static_cast<int>(tid - a * GS);
  const bool on = a < n_groups && s < S;
  const int64_t aa = a < n_groups ? a : 0;  // threads past the end read node 0, then drop it
  const int ss = s < S ? s : 0;
  const int64_t da = gdof[aa];
  int64_t db[EM];
#pragma unroll
  for (int k = 0; k < EM; ++k) {
    const int32_t b = gnb[static_cast<int64_t>(k) * n_groups + aa];
    db[k] = b >= 0 ? static_cast<int64_t>(gdof[b]) : da;
  }
  double ua1[GD], ua2[GD], ub1[EM][GD];
#pragma unroll
  for (int c = 0; c < GD; ++c) {
    ua1[c] = static_cast<double>(um1[(da + c) * S + ss]);
    ua2[c] = static_cast<double>(um2[(da + c) * S + ss]);
  }
#pragma unroll
  for (int k = 0; k < EM; ++k)
#pragma unroll
    for (int c = 0; c < GD; ++c) ub1[k][c] = static_cast<double>(um1[(db[k] + c) * S + ss]);
  double cs[BS], d0[GD], d2[GD], cab[EM][BS], cba[EM][BS];
#pragma unroll
  for (int i = 0; i < BS; ++i) cs[i] = 0.0;
#pragma unroll
  for (int c = 0; c < GD; ++c) d0[c] = d2[c] = 0.0;
#pragma unroll
  for (int k = 0; k < EM; ++k)
#pragma unroll
    for (int i = 0; i < BS; ++i) cab[k][i] = cba[k][i] = 0.0;
  for (int32_t t = 0; t < Ls; ++t) {
    double ua[GD], la[GD], ub[EM][GD], lb[EM][GD];
#pragma unroll
    for (int c = 0; c < GD; ++c) {
      ua[c] = static_cast<double>(U[t * us + (da + c) * ud + ss]);
      la[c] = static_cast<double>(Lam[t * ls + (da + c) * ld + ss]);
    }
#pragma unroll
    for (int k = 0; k < EM; ++k)
#pragma unroll
      for (int c = 0; c < GD; ++c) {
        ub[k][c] = static_cast<double>(U[t * us + (db[k] + c) * ud + ss]);
        lb[k][c] = static_cast<double>(Lam[t * ls + (db[k] + c) * ld + ss]);
      }
#pragma unroll
    for (int c = 0; c < GD; ++c) {
      d0[c] += la[c] * ua[c];
      d2[c] += la[c] * ua2[c];
#pragma unroll
      for (int d = 0; d < GD; ++d) cs[c * GD + d] += la[c] * ua1[d];
    }
#pragma unroll
    for (int k = 0; k < EM; ++k)
#pragma unroll
      for (int c = 0; c < GD; ++c)
#pragma unroll
        for (int d = 0; d < GD; ++d) {
          cab[k][c * GD + d] += la[c] * ub1[k][d];
          cba[k][c * GD + d] += lb[k][c] * ua1[d];
        }
#pragma unroll
    for (int c = 0; c < GD; ++c) {
      ua2[c] = ua1[c];
      ua1[c] = ua[c];
    }
#pragma unroll
    for (int k = 0; k < EM; ++k)
#pragma unroll
      for (int c = 0; c < GD; ++c) ub1[k][c] = ub[k][c];
  }
  // the node's record: self block, D0, D2, then (C[a,b], C[b,a]) per slot
  double r[STRIDE];
#pragma unroll
  for (int i = 0; i < BS; ++i) r[i] = cs[i];
#pragma unroll
  for (int c = 0; c < GD; ++c) {
    r[BS + c] = d0[c];
    r[BS + GD + c] = d2[c];
  }
#pragma unroll
  for (int k = 0; k < EM; ++k)
#pragma unroll
    for (int i = 0; i < BS; ++i) {
      r[BS + 2 * GD + k * 2 * BS + i] = cab[k][i];
      r[BS + 2 * GD + k * 2 * BS + BS + i] = cba[k][i];
    }
#pragma unroll
  for (int i = 0; i < STRIDE; ++i) {
    double x = on ? r[i] : 0.0;
    for (int o = GS >> 1; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);  // the node's lanes
    r[i] = x;
  }
  if (s == 0 && a < n_groups) {
    double* out = acc + a * STRIDE;
#pragma unroll
    for (int i = 0; i < STRIDE; ++i) out[i] += r[i];
  }
}

// Element gradients from the node-pair records: eoff[e, i, j] is the offset
// of the GD x GD block C[vertex i's node, vertex j's node] in acc (-1 where a
// vertex is absent), egrp[e, i] the node of vertex i.
//   gK_e -= dt^2 sum_{i,j,c,d} K_e[i GD + c, j GD + d] C_ij[c, d]
//   gM_e -= sum_{i,c} m_e[i GD + c] (w0 D0 + w1 C[a, a]_cc + w2 D2)[c]
template <typename T, int KV, int GD>
__global__ __launch_bounds__(kBlock) void k_pair_contract(
    int32_t n_el, int32_t stride, const int32_t* __restrict__ egrp,
    const int64_t* __restrict__ eoff, const T* __restrict__ el_K, const T* __restrict__ el_m,
    const T* __restrict__ el_w, const double* __restrict__ acc, double dt2,
    double* __restrict__ gK, double* __restrict__ gM) {
  constexpr int KM = KV * GD, BS = GD * GD;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_el) return;
  double sk = 0.0, sm = 0.0;
#pragma unroll
  for (int i = 0; i < KV; ++i)
#pragma unroll
    for (int j = 0; j < KV; ++j) {
      const int64_t off = eoff[(e * KV + i) * KV + j];
      if (off < 0) continue;
#pragma unroll
      for (int c = 0; c < GD; ++c)
#pragma unroll
        for (int d = 0; d < GD; ++d)
          sk += static_cast<double>(el_K[(e * KM + i * GD + c) * KM + j * GD + d]) *
                acc[off + c * GD + d];
    }
  const double w0 = static_cast<double>(el_w[3 * e]), w1 = static_cast<double>(el_w[3 * e + 1]),
               w2 = static_cast<double>(el_w[3 * e + 2]);
#pragma unroll
  for (int i = 0; i < KV; ++i) {
    const int32_t a = egrp[e * KV + i];
    if (a < 0) continue;
    const double* r = acc + static_cast<int64_t>(a) * stride;
#pragma unroll
    for (int c = 0; c < GD; ++c)
      sm += static_cast<double>(el_m[e * KM + i * GD + c]) *
            (w0 * r[BS + c] + w1 * r[c * GD + c] + w2 * r[BS + GD + c]);
  }
  gK[e] -= dt2 * sk;
  gM[e] -= sm;
}

// ---------------------------------------------------------------------------
// Device buffers owned by a handle or a run.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) : bytes(b) {
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw std::bad_alloc();
    }
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename U> U* as() const { return static_cast<U*>(p); }
};

// Blocking copy: the source is often a temporary, and the data must be on
// the device before any kernel of the handle's stream reads it.
template <typename U>
DevBuf upload(const std::vector<U>& v) {
  DevBuf b(v.size() * sizeof(U));
  if (!v.empty()) BFE_HIP(hipMemcpy(b.p, v.data(), b.bytes, hipMemcpyHostToDevice));
  return b;
}

template <typename T>
std::vector<T> cast_vec(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

// Groups (row, payload) entries by block of `rows_per_block` rows.
struct Grouped {
  std::vector<int32_t> ptr, row, a, b;
};

Grouped group_by_block(int64_t nblk, int64_t rows_per_block, const std::vector<int32_t>& rows,
                       const std::vector<int32_t>* a, const std::vector<int32_t>* b) {
  Grouped G;
  G.ptr.assign(nblk + 1, 0);
  for (int32_t r : rows) G.ptr[r / rows_per_block + 1]++;
  for (int64_t k = 0; k < nblk; ++k) G.ptr[k + 1] += G.ptr[k];
  const size_t m = rows.size();
  G.row.resize(m);
  if (a) G.a.resize(m);
  if (b) G.b.resize(m);
  std::vector<int32_t> fill(G.ptr.begin(), G.ptr.end() - 1);
  for (size_t i = 0; i < m; ++i) {
    const int32_t o = fill[rows[i] / rows_per_block]++;
    G.row[o] = rows[i];
    if (a) G.a[o] = (*a)[i];
    if (b) G.b[o] = (*b)[i];
  }
  return G;
}

// One group of the element-by-element operator: elements of one type over
// the group's own node numbering (k_step_ebe / k_step_ebe_pat).
struct EbeGroup {
  int64_t n_nodes = 0, n_el = 0;
  int32_t dpn = 0, npe = 0, n_slot = 0, n_tmpl = 0, n_cls = 0;
  int32_t pattern = 0;       // mesh pattern of k_step_ebe_pat, 0: none
  bool uniform = false;      // node j's first DOF is j*dpn (no DOF table read)
  bool smask_all = false;    // most occupied slots scaled (anchor table, no smask)
  bool sparse = false;       // slot rows zero (to rounding residue) outside PatKuhn24::nz
  int32_t coef_node = 0;     // alpha/beta per group node (alpha_n, beta_n)
  std::vector<int64_t> doff; // first-DOF offset of each neighbour class on regular nodes
  std::vector<int32_t> anc_noff, anc_row;  // anchor table column offset and row per slot
  DevBuf dof0, adj, nodes, scale, srows, nbr, srows_c, pmask, smask, stencil, anchor;
  DevBuf alpha_n, beta_n;    // per node (coef_node)
  DevBuf iptr, irow, ik;     // interface rows (group-local rows) per workgroup
  DevBuf nblk;               // node-block form (built at the first batched run that uses it)
  int64_t bytes() const {
    int64_t s = 0;
    for (const DevBuf* b : {&dof0, &adj, &nodes, &scale, &srows, &nbr, &srows_c, &pmask, &smask,
                            &stencil, &anchor, &alpha_n, &beta_n, &iptr, &irow, &ik, &nblk})
      s += static_cast<int64_t>(b->bytes);
    return s;
  }
};

// Element-by-element operator attached to a handle: K = sum over groups.
struct EbeOp {
  std::vector<EbeGroup> groups;
  // global row -> (group, group-local row); group -1 when no group owns it
  std::vector<int8_t> row_group;
  std::vector<int32_t> row_local;
  DevBuf alpha, beta;         // per DOF, for groups without per-node coefficients
  double check_err = 0.0;     // largest relative difference to the CSR step at set-up
  int64_t n_el() const {
    int64_t s = 0;
    for (const auto& g : groups) s += g.n_el;
    return s;
  }
  int64_t bytes() const {
    int64_t s = static_cast<int64_t>(alpha.bytes + beta.bytes);
    for (const auto& g : groups) s += g.bytes();
    return s;
  }
  // alpha/beta a group's kernel reads: per group node or per global DOF
  template <typename T>
  const T* alpha_of(const EbeGroup& g) const {
    return g.coef_node ? g.alpha_n.as<T>() : alpha.as<T>();
  }
  template <typename T>
  const T* beta_of(const EbeGroup& g) const {
    return g.coef_node ? g.beta_n.as<T>() : beta.as<T>();
  }
};

}  // namespace

namespace bfe {
// The subdomain operations (bfe_subdom.hip) report through bfe_last_error.
void set_subdomain_error(const std::string& msg) { g_last_error = msg; }
}  // namespace bfe

// An operator of the multi-source march: row groups over column unions, rows
// contiguous within each 64-row slice (see group_op).
struct GroupOp {
  DevBuf off, col, val;
  int64_t gathers = 0;  // union entries summed over the groups
  bool built = false;
};

struct bfe_plan_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int64_t n = 0;
  int mode = 0, prec = 0;
  double dt = 0.0;
  bool damped = false;
  int64_t n_slices = 0, nnz_sell = 0, nnz_k = 0;
  int32_t nI = 0;
  int64_t nnz_coff = 0, nnz_xinv = 0;
  bool has_xinv = false;
  double sq = 0.0, sr = 0.0;
  std::vector<double> alpha;
  DevBuf off, col, val, beta;
  DevBuf iface, cptr, ccol, cval, gfix, xptr, xcol, xval;
  // interface-row lists grouped per workgroup of the single-source march
  DevBuf is_ptr, is_row, is_k;
  bfe_run_stats stats{};
  int32_t timing_every = 0;
  int32_t multi_rows = 2;   // wavefronts per workgroup of k_step_multi
  int32_t multi_rpw = 4;    // rows per wavefront of k_step_multi
  int32_t multi_group = 0;  // rows per gather group of k_step_multi (0: automatic)
  int32_t auto_group = 0;   // automatic choice once made (1, 2 or 4)
  int32_t last_group = 0;   // group of the last batched run
  GroupOp grp[4];           // multi-source operators for G = 1, 2, 4, 8 (built on use)
  int32_t xcd_remap = 1;    // XCD-aware workgroup order
  std::vector<hipEvent_t> tev;  // pairs (start, stop) for sampled launches
  // adjoint support (defined further down)
  std::map<int32_t, std::shared_ptr<void>> adj_records;  // AdjointRecord per slot
  std::shared_ptr<void> grad_elems;  // GradElements
  std::vector<int32_t> iface_host;   // interface rows (global), for regrouping
  std::vector<double> beta_host;     // beta per row (damped plans), for the element operator
  std::vector<double> m_host;        // lumped mass per DOF (field output)
  std::shared_ptr<void> field_elems; // FieldElements
  std::unique_ptr<EbeOp> ebe;        // optional element-by-element operator
  int32_t use_ebe = 1;               // single-source march uses ebe when set
  int32_t ebe_group = 8;             // slots per load group of k_step_ebe (2, 4 or 8)
  int32_t ebe_kernel = 1;            // 1: pattern kernel when the mesh matches one
  int32_t use_graph = 0;             // single-source march captured in a hipGraph
  int32_t use_persistent = 0;        // single-source CSR march as one cooperative launch
  int32_t persistent_grid = 0;       // cap on its workgroups (0: as many as fit at once)
  int32_t multi_ebe = 0;             // batched march's operator: BFE_OPERATOR_CSR, _EBE or _NODE
  int32_t multi_npw = 4;             // nodes per wavefront of k_step_ebe_multi
  int32_t multi_pipeline = 1;        // k_step_ebe_multi: neighbour rows slide along node runs
  int32_t grad_pairs = 1;            // batched adjoint: node-pair gradients where possible
  int32_t last_multi_op = BFE_OPERATOR_CSR;  // operator of the last batched run
  ~bfe_plan_s() {
    if (stream) {
      (void)hipSetDevice(device);
      (void)hipStreamSynchronize(stream);
    }
    adj_records.clear();
    grad_elems.reset();
    field_elems.reset();
    ebe.reset();
    for (DevBuf* b : {&off, &col, &val, &beta, &iface, &cptr, &ccol, &cval, &gfix, &xptr, &xcol,
                      &xval, &is_ptr, &is_row, &is_k})
      b->release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int64_t device_bytes() const {
    int64_t s = 0;
    for (const DevBuf* b : {&off, &col, &val, &beta, &iface, &cptr, &ccol, &cval, &gfix, &xptr,
                            &xcol, &xval, &is_ptr, &is_row, &is_k})
      s += static_cast<int64_t>(b->bytes);
    for (const GroupOp& g : grp) s += static_cast<int64_t>(g.off.bytes + g.col.bytes + g.val.bytes);
    if (ebe) s += ebe->bytes();
    return s;
  }
};

namespace {

template <typename F>
bfe_status guarded(F&& fn) {
  try {
    fn();
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    g_last_error = std::string("invalid argument: ") + e.what();
    return BFE_ERR_ARG;
  } catch (const std::domain_error& e) {
    g_last_error = std::string("numerical error: ") + e.what();
    return BFE_ERR_NUMERIC;
  } catch (const hip_failure& e) {
    g_last_error = std::string("HIP error: ") + e.what();
    return BFE_ERR_HIP;
  } catch (const std::bad_alloc&) {
    g_last_error = "out of memory";
    return BFE_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = std::string("internal error: ") + e.what();
    return BFE_ERR_INTERNAL;
  } catch (...) {
    g_last_error = "internal error";
    return BFE_ERR_INTERNAL;
  }
}

// Step-kernel sampling: event pair k brackets the k-th sampled launch.
struct KernelSampler {
  bfe_plan_s& H;
  int64_t t0;
  int64_t used = 0;
  KernelSampler(bfe_plan_s& h, int64_t first_step, int64_t n_steps) : H(h), t0(first_step) {
    if (H.timing_every <= 0) return;
    const int64_t need = 2 * ((n_steps + H.timing_every - 1) / H.timing_every);
    while (static_cast<int64_t>(H.tev.size()) < need) {
      hipEvent_t e;
      BFE_HIP(hipEventCreate(&e));
      H.tev.push_back(e);
    }
  }
  bool sampled(int64_t t) const { return H.timing_every > 0 && (t - t0) % H.timing_every == 0; }
  void before(int64_t t) {
    if (sampled(t)) BFE_HIP(hipEventRecord(H.tev[2 * used], H.stream));
  }
  void after(int64_t t) {
    if (sampled(t)) {
      BFE_HIP(hipEventRecord(H.tev[2 * used + 1], H.stream));
      ++used;
    }
  }
  // call after the stream has been synchronised
  void collect() {
    double total = 0.0;
    for (int64_t k = 0; k < used; ++k) {
      float ms = 0.f;
      BFE_HIP(hipEventElapsedTime(&ms, H.tev[2 * k], H.tev[2 * k + 1]));
      total += ms;
    }
    H.stats.step_kernel_ms = total;
    H.stats.step_kernel_samples = used;
  }
};

template <typename T>
void upload_plan(bfe_plan_s& H, const bfe::Plan& P) {
  if (!P.op.off.empty()) {  // SELL built on the host
    H.off = upload(P.op.off);
    H.col = upload(P.op.col);
    H.val = upload(cast_vec<T>(P.op.val));
  }
  if (P.damped) H.beta = upload(cast_vec<T>(P.beta));
  const int64_t n = P.n;
  const int64_t nb_s = (n + kRowsSingle - 1) / kRowsSingle;
  std::vector<int32_t> ks(P.iface.size());
  for (size_t k = 0; k < ks.size(); ++k) ks[k] = static_cast<int32_t>(k);
  Grouped gs = group_by_block(nb_s, kRowsSingle, P.iface, &ks, nullptr);
  H.is_ptr = upload(gs.ptr); H.is_row = upload(gs.row); H.is_k = upload(gs.a);
  if (!P.iface.empty()) {
    H.iface = upload(P.iface);
    H.cptr = upload(P.coff.ptr);
    H.ccol = upload(P.coff.col);
    H.cval = upload(cast_vec<T>(P.coff.val));
    H.gfix = upload(cast_vec<T>(P.gfix));
    if (P.has_xinv) {
      H.xptr = upload(P.xinv.ptr);
      H.xcol = upload(P.xinv.col);
      H.xval = upload(cast_vec<T>(P.xinv.val));
    }
  }
}

// SELL-64 of alpha-scaled K built on the device from K's CSR (the layout of
// bfe_plan.cpp's host builder, value for value): slice widths, then one
// thread per (slice, lane) writes its row's entries and padding.
__global__ void k_sell_width(int64_t n, int64_t n_slices, const int64_t* __restrict__ kp,
                             int64_t* __restrict__ width) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n_slices) return;
  int64_t w = 0;
  for (int64_t r = s * 64; r < min(n, s * 64 + 64); ++r) w = max(w, kp[r + 1] - kp[r]);
  width[s] = w;
}

template <typename T>
__global__ void k_sell_fill(int64_t n, int64_t n_slices, const int64_t* __restrict__ kp,
                            const int32_t* __restrict__ kc, const double* __restrict__ kv,
                            const double* __restrict__ alpha, const int64_t* __restrict__ off,
                            int32_t* __restrict__ col, T* __restrict__ val) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t s = r >> 6;
  if (s >= n_slices) return;
  const int lane = static_cast<int>(r & 63);
  const int64_t w = (off[s + 1] - off[s]) / 64;
  const int64_t b = r < n ? kp[r] : 0, e = r < n ? kp[r + 1] : 0;
  const int32_t pad = static_cast<int32_t>(min(r, n - 1));
  for (int64_t k = 0; k < w; ++k) {
    const int64_t o = off[s] + 64 * k + lane;
    if (b + k < e) {
      col[o] = kc[b + k];
      val[o] = static_cast<T>(alpha[r] * kv[b + k]);
    } else {
      col[o] = pad;
      val[o] = T(0);
    }
  }
}

// SELL-64 of alpha-scaled K, built on the device into (off, col, val);
// returns the stored entry count including padding.
template <typename T>
int64_t build_sell_device(bfe_plan_s& H, const bfe::CsrRef& K, const std::vector<double>& alpha,
                          DevBuf& off_out, DevBuf& col_out, DevBuf& val_out) {
  const int64_t n = K.n, ns = (n + 63) / 64, nnz = K.nnz();
  DevBuf kp = upload(std::vector<int64_t>(K.ptr, K.ptr + n + 1));
  DevBuf kc(sizeof(int32_t) * std::max<int64_t>(nnz, 1)), kv(sizeof(double) * std::max<int64_t>(nnz, 1));
  if (nnz > 0) {
    BFE_HIP(hipMemcpy(kc.p, K.col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    BFE_HIP(hipMemcpy(kv.p, K.val, sizeof(double) * nnz, hipMemcpyHostToDevice));
  }
  DevBuf width(sizeof(int64_t) * ns);
  hipLaunchKernelGGL(k_sell_width, dim3((ns + 255) / 256), dim3(256), 0, H.stream, n, ns,
                     kp.as<int64_t>(), width.as<int64_t>());
  BFE_HIP(hipGetLastError());
  std::vector<int64_t> off(ns + 1, 0);
  BFE_HIP(hipMemcpyAsync(off.data() + 1, width.p, sizeof(int64_t) * ns, hipMemcpyDeviceToHost,
                         H.stream));
  BFE_HIP(hipStreamSynchronize(H.stream));
  for (int64_t q = 0; q < ns; ++q) off[q + 1] = off[q] + 64 * off[q + 1];
  DevBuf d_alpha = upload(alpha);
  off_out = upload(off);
  col_out = DevBuf(sizeof(int32_t) * std::max<int64_t>(off[ns], 1));
  val_out = DevBuf(sizeof(T) * std::max<int64_t>(off[ns], 1));
  hipLaunchKernelGGL((k_sell_fill<T>), dim3((ns * 64 + 255) / 256), dim3(256), 0, H.stream, n,
                     ns, kp.as<int64_t>(), kc.as<int32_t>(), kv.as<double>(),
                     d_alpha.as<double>(), off_out.as<int64_t>(), col_out.as<int32_t>(),
                     val_out.as<T>());
  BFE_HIP(hipGetLastError());
  BFE_HIP(hipStreamSynchronize(H.stream));
  return off[ns];
}

template <typename T>
void build_operators(bfe_plan_s& H, const bfe::CsrRef& K, const std::vector<double>& alpha) {
  H.n_slices = (K.n + 63) / 64;
  H.nnz_sell = build_sell_device<T>(H, K, alpha, H.off, H.col, H.val);
}

// ---------------------------------------------------------------------------
// Operators of the multi-source march, built on the device from the SELL-64
// operator: each aligned group of G rows (G = 1, 2, 4 or 8; G divides 64, so
// a group lies in one slice) stored over the union of its rows' columns,
// explicit zeros where a row lacks a column, padded per slice to the widest
// union with value 0. Rows are contiguous within a slice (entry k of row r
// at off[s] + (r % 64) w_s + k, w_s the slice width), so the entry loads of
// a wavefront's rows are coalesced. Entries of value 0 in the source
// (padding) are skipped.
template <typename T, int G, typename Emit>
__device__ __forceinline__ int group_merge(int64_t base, int w, const int32_t* __restrict__ col,
                                           const T* __restrict__ val, Emit&& emit) {
  int h[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    h[j] = 0;
    while (h[j] < w && val[base + j + 64 * static_cast<int64_t>(h[j])] == T(0)) ++h[j];
  }
  int len = 0;
  for (;;) {
    int c = INT32_MAX;
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (h[j] < w) c = min(c, col[base + j + 64 * static_cast<int64_t>(h[j])]);
    if (c == INT32_MAX) break;
    T v[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      v[j] = T(0);
      if (h[j] < w && col[base + j + 64 * static_cast<int64_t>(h[j])] == c) {
        v[j] = val[base + j + 64 * static_cast<int64_t>(h[j])];
        do ++h[j];
        while (h[j] < w && val[base + j + 64 * static_cast<int64_t>(h[j])] == T(0));
      }
    }
    emit(len, c, v);
    ++len;
  }
  return len;
}

template <typename T, int G>
__global__ void k_group_width(int64_t n_slices, const int64_t* __restrict__ off,
                              const int32_t* __restrict__ col, const T* __restrict__ val,
                              unsigned long long* __restrict__ width,
                              unsigned long long* __restrict__ total) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t s = g / (64 / G);
  if (s >= n_slices) return;
  const int r0 = static_cast<int>(g % (64 / G)) * G;  // first row of the group in its slice
  const int w = static_cast<int>((off[s + 1] - off[s]) >> 6);
  const int len = group_merge<T, G>(off[s] + r0, w, col, val, [](int, int, const T*) {});
  atomicMax(width + s, static_cast<unsigned long long>(len));
  atomicAdd(total, static_cast<unsigned long long>(len));
}

template <typename T, int G>
__global__ void k_group_fill(int64_t n, int64_t n_slices, const int64_t* __restrict__ off,
                             const int32_t* __restrict__ col, const T* __restrict__ val,
                             const int64_t* __restrict__ goff, int32_t* __restrict__ gcol,
                             T* __restrict__ gval) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t s = g / (64 / G);
  if (s >= n_slices) return;
  const int r0 = static_cast<int>(g % (64 / G)) * G;
  const int w = static_cast<int>((off[s + 1] - off[s]) >> 6);
  const int wg = static_cast<int>((goff[s + 1] - goff[s]) >> 6);
  const int64_t ob = goff[s] + static_cast<int64_t>(r0) * wg;  // row j at ob + j wg
  const int len = group_merge<T, G>(off[s] + r0, w, col, val, [&](int k, int c, const T* v) {
#pragma unroll
    for (int j = 0; j < G; ++j) {
      gcol[ob + static_cast<int64_t>(j) * wg + k] = c;
      gval[ob + static_cast<int64_t>(j) * wg + k] = v[j];
    }
  });
  for (int k = len; k < wg; ++k)
#pragma unroll
    for (int j = 0; j < G; ++j) {
      gcol[ob + static_cast<int64_t>(j) * wg + k] =
          static_cast<int32_t>(min<int64_t>(s * 64 + r0 + j, n - 1));
      gval[ob + static_cast<int64_t>(j) * wg + k] = T(0);
    }
}

int group_index(int G) { return G == 1 ? 0 : G == 2 ? 1 : G == 4 ? 2 : 3; }

template <typename T, int G>
void build_group_op_g(bfe_plan_s& H, GroupOp& op) {
  const int64_t ns = H.n_slices, ng = ns * (64 / G);
  DevBuf width(sizeof(unsigned long long) * ns), total(sizeof(unsigned long long));
  BFE_HIP(hipMemsetAsync(width.p, 0, width.bytes, H.stream));
  BFE_HIP(hipMemsetAsync(total.p, 0, total.bytes, H.stream));
  hipLaunchKernelGGL((k_group_width<T, G>), dim3((ng + 255) / 256), dim3(256), 0, H.stream, ns,
                     H.off.as<int64_t>(), H.col.as<int32_t>(), H.val.as<T>(),
                     width.as<unsigned long long>(), total.as<unsigned long long>());
  BFE_HIP(hipGetLastError());
  std::vector<int64_t> off(ns + 1, 0);
  unsigned long long tot = 0;
  BFE_HIP(hipMemcpyAsync(off.data() + 1, width.p, width.bytes, hipMemcpyDeviceToHost, H.stream));
  BFE_HIP(hipMemcpyAsync(&tot, total.p, sizeof(tot), hipMemcpyDeviceToHost, H.stream));
  BFE_HIP(hipStreamSynchronize(H.stream));
  for (int64_t q = 0; q < ns; ++q) off[q + 1] = off[q] + 64 * off[q + 1];
  op.off = upload(off);
  op.col = DevBuf(sizeof(int32_t) * std::max<int64_t>(off[ns], 1));
  op.val = DevBuf(sizeof(T) * std::max<int64_t>(off[ns], 1));
  hipLaunchKernelGGL((k_group_fill<T, G>), dim3((ng + 255) / 256), dim3(256), 0, H.stream, H.n,
                     ns, H.off.as<int64_t>(), H.col.as<int32_t>(), H.val.as<T>(),
                     op.off.as<int64_t>(), op.col.as<int32_t>(), op.val.as<T>());
  BFE_HIP(hipGetLastError());
  BFE_HIP(hipStreamSynchronize(H.stream));
  op.gathers = static_cast<int64_t>(tot);
  op.built = true;
}

// The operator for G (1, 2, 4 or 8), built on first use.
template <typename T>
const GroupOp& group_op(bfe_plan_s& H, int G) {
  GroupOp& op = H.grp[group_index(G)];
  if (!op.built) {
    if (G == 1) build_group_op_g<T, 1>(H, op);
    else if (G == 2) build_group_op_g<T, 2>(H, op);
    else if (G == 4) build_group_op_g<T, 4>(H, op);
    else build_group_op_g<T, 8>(H, op);
  }
  return op;
}

// Automatic row groups: pairs when they save at least a third of the
// gathers (2D solids: about 40 %), and groups of 4 (two neighbouring nodes)
// when those save a further fifth of the pairs' gathers (2D solids: 28 %;
// 2.0 against 2.3 ms per 64-column step on 4.9M DOF); else one gather per
// row. The operators not chosen are released.
constexpr double kPairMaxRatio = 0.65;  // pair gathers / gathers of K
constexpr double kQuadMaxRatio = 0.8;   // group-of-4 gathers / pair gathers

template <typename T>
int resolve_group(bfe_plan_s& H, int rpw) {
  int G = H.multi_group;
  if (G == 0) {
    if (H.auto_group == 0) {
      const double g2 = static_cast<double>(group_op<T>(H, 2).gathers);
      if (g2 <= kPairMaxRatio * static_cast<double>(H.nnz_k)) {
        const double g4 = static_cast<double>(group_op<T>(H, 4).gathers);
        H.auto_group = g4 <= kQuadMaxRatio * g2 ? 4 : 2;
      } else {
        H.auto_group = 1;
      }
      for (int g : {2, 4})
        if (g != H.auto_group) H.grp[group_index(g)] = GroupOp{};
    }
    G = H.auto_group;
  }
  G = std::min(G, rpw);
  group_op<T>(H, G);
  return G;
}

// 0-based first step of the march, -1 if none: v6:163-166 takes
// min(find(sum(abs(forcing_functions)))). With several rows the sum runs down
// each column, so the march starts at the first column with a non-zero (or
// NaN) entry. A single row (one signal broadcast to every forcing DOF) sums
// to one scalar, so the march starts at step 0 whenever any sample is
// non-zero, leading zero samples included.
int64_t first_forced_step(const double* F, int32_t rows, int32_t n_t) {
  if (rows == 1) {
    for (int64_t t = 0; t < n_t; ++t)
      if (F[t] != 0.0) return 0;  // NaN compares unequal to 0 as in MATLAB find()
    return -1;
  }
  for (int64_t t = 0; t < n_t; ++t) {
    double s = 0.0;
    for (int32_t r = 0; r < rows; ++r) s += std::fabs(F[t * rows + r]);
    if (s != 0.0) return t;  // NaN compares unequal to 0 as in MATLAB find()
  }
  return -1;
}

void check_indices(const int32_t* idx, int64_t count, int64_t n, const char* what) {
  if (count > 0 && idx == nullptr) throw std::invalid_argument(std::string(what) + " is NULL");
  for (int64_t i = 0; i < count; ++i)
    if (idx[i] < 0 || idx[i] >= n)
      throw std::invalid_argument(std::string(what) + " index out of range");
}

// Everything a run needs on the device besides the plan.
struct RunBuffers {
  DevBuf F, fptr, frow, fsrc, ffrow, falpha, mptr, mrow, mk, hidx, wI;
  BlockLists lists(const bfe_plan_s& H) const {
    BlockLists L;
    L.fptr = fptr.as<int32_t>(); L.f_row = frow.as<int32_t>(); L.f_src = fsrc.as<int32_t>();
    L.f_frow = ffrow.as<int32_t>();
    L.mptr = mptr.as<int32_t>(); L.m_row = mrow.as<int32_t>(); L.m_k = mk.as<int32_t>();
    L.iptr = H.is_ptr.as<int32_t>();
    L.i_row = H.is_row.as<int32_t>();
    L.i_k = H.is_k.as<int32_t>();
    return L;
  }
};

// Builds forcing and monitor lists. Forcing entries are (dof, source
// column, signal row); for one (dof, source) the last entry wins, like
// MATLAB's f(idx) = v. Forcing on an interface row enters w_I with
// alpha = dt^2/m, exactly as f enters y before X_II^-1 in v6:136.
template <typename T>
void make_run_buffers(RunBuffers& R, const bfe_plan_s& H, int64_t rows_per_block,
                      int64_t nblk, int32_t S, const int32_t* frc_idx, const int32_t* frc_src,
                      int32_t p, const double* frcs, int32_t frc_rows, int32_t n_t,
                      const int32_t* hist_idx, int32_t q) {
  std::vector<std::pair<std::pair<int32_t, int32_t>, int32_t>> key;
  for (int32_t i = 0; i < p; ++i) key.push_back({{frc_idx[i], frc_src ? frc_src[i] : 0}, i});
  std::stable_sort(key.begin(), key.end(), [](auto& a, auto& b) { return a.first < b.first; });
  std::vector<int32_t> rows, srcs, frow;
  for (size_t i = 0; i < key.size(); ++i)
    if (i + 1 == key.size() || key[i + 1].first != key[i].first) {
      rows.push_back(key[i].first.first);
      srcs.push_back(key[i].first.second);
      frow.push_back(frc_rows == 1 ? 0 : key[i].second);
    }
  Grouped gf = group_by_block(nblk, rows_per_block, rows, &srcs, &frow);
  std::vector<double> fa(gf.row.size());
  for (size_t i = 0; i < fa.size(); ++i) fa[i] = H.alpha[gf.row[i]];
  std::vector<int32_t> mrows(hist_idx, hist_idx + q), mks(q);
  for (int32_t k = 0; k < q; ++k) mks[k] = k;
  Grouped gmon = group_by_block(nblk, rows_per_block, mrows, &mks, nullptr);
  std::vector<double> Fh(frcs, frcs + static_cast<size_t>(frc_rows) * n_t);
  R.F = upload(cast_vec<T>(Fh));
  R.fptr = upload(gf.ptr); R.frow = upload(gf.row); R.fsrc = upload(gf.a);
  R.ffrow = upload(gf.b); R.falpha = upload(cast_vec<T>(fa));
  R.mptr = upload(gmon.ptr); R.mrow = upload(gmon.row); R.mk = upload(gmon.a);
  R.hidx = upload(mrows);
  R.wI = DevBuf(sizeof(T) * static_cast<size_t>(H.nI) * S);
}

template <typename T>
IfaceRows<T> iface_rows(const bfe_plan_s& H, T* wI) {
  IfaceRows<T> I;
  I.cptr = H.cptr.as<int64_t>();
  I.ccol = H.ccol.as<int32_t>();
  I.cval = H.cval.as<T>();
  I.gfix = H.gfix.as<T>();
  I.sq = static_cast<T>(H.sq);
  I.sr = static_cast<T>(H.sr);
  I.wI = H.has_xinv ? wI : nullptr;
  return I;
}

template <typename T>
void launch_iface(const bfe_plan_s& H, int32_t S, const T* wI, T* u) {
  if (S == 1) {
    const int64_t work = static_cast<int64_t>(H.nI) * kIfaceLanes;
    hipLaunchKernelGGL((k_iface<T>), dim3((work + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       H.stream, H.nI, H.iface.as<int32_t>(), H.xptr.as<int64_t>(),
                       H.xcol.as<int32_t>(), H.xval.as<T>(), wI, u);
  } else {
    const int64_t work = static_cast<int64_t>(H.nI) * S;
    hipLaunchKernelGGL((k_iface_multi<T>), dim3((work + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       H.stream, H.nI, S, H.iface.as<int32_t>(), H.xptr.as<int64_t>(),
                       H.xcol.as<int32_t>(), H.xval.as<T>(), wI, u);
  }
}

// One cooperative launch of k_march_persistent for steps t0 .. n_t - 1, with
// as many workgroups as the step has blocks, at most as many as the device
// holds at once (cooperative launch refuses more).
template <typename T>
void launch_persistent(const bfe_plan_s& H, int64_t nblk, int64_t t0, int32_t n_t, T* ubuf,
                       const BlockLists& L, const T* falpha, const T* F, int32_t F_rows,
                       int32_t q, T* hist, int32_t field_every, T* fld, int32_t fld_is_vel, T dt,
                       const IfaceRows<T>& IR) {
  PersistentMarch<T> a{};
  a.n = H.n;
  a.nblk = nblk;
  a.t0 = t0;
  a.n_t = n_t;
  a.off = H.off.as<int64_t>();
  a.col = H.col.as<int32_t>();
  a.val = H.val.as<T>();
  a.beta = H.damped ? H.beta.as<T>() : nullptr;
  a.ubuf = ubuf;
  a.L = L;
  a.f_alpha = falpha;
  a.F = F;
  a.F_rows = F_rows;
  a.q = q;
  a.hist = hist;
  a.field_every = field_every;
  a.fld_is_vel = fld_is_vel;
  a.fld = fld;
  a.dt = dt;
  a.I = IR;
  a.nI = (H.nI > 0 && H.has_xinv) ? H.nI : 0;
  a.iface = H.iface.as<int32_t>();
  a.xptr = H.xptr.as<int64_t>();
  a.xcol = H.xcol.as<int32_t>();
  a.xval = H.xval.as<T>();
  const bool nt = H.nnz_sell * static_cast<int64_t>(sizeof(T) + sizeof(int32_t)) > kStreamBytes;
  const void* fn =
      H.damped ? (nt ? reinterpret_cast<const void*>(&k_march_persistent<T, true, true>)
                     : reinterpret_cast<const void*>(&k_march_persistent<T, true, false>))
               : (nt ? reinterpret_cast<const void*>(&k_march_persistent<T, false, true>)
                     : reinterpret_cast<const void*>(&k_march_persistent<T, false, false>));
  int coop = 0, bpcu = 0, ncu = 0;
  BFE_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, H.device));
  if (!coop) throw hip_failure("persistent march: the device has no cooperative launch");
  BFE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpcu, fn, kBlock, 0));
  BFE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, H.device));
  const int64_t cap = static_cast<int64_t>(bpcu) * ncu;
  if (cap < 1) throw hip_failure("persistent march: no co-resident workgroups");
  int64_t grid = std::min(nblk, cap);
  if (H.persistent_grid > 0) grid = std::min<int64_t>(grid, H.persistent_grid);
  void* args[] = {&a};
  BFE_HIP(hipLaunchCooperativeKernel(fn, dim3(static_cast<uint32_t>(grid)), dim3(kBlock), args, 0,
                                     H.stream));
}

void validate_run(int64_t n, int32_t n_t, const int32_t* frc_idx, int32_t p, const double* frcs,
                  int32_t frc_rows, const int32_t* hist_idx, int32_t q, const double* hist_out) {
  if (n_t < 1) throw std::invalid_argument("n_t must be >= 1");
  if (p < 0 || q < 0) throw std::invalid_argument("negative count");
  check_indices(frc_idx, p, n, "forcing");
  check_indices(hist_idx, q, n, "history");
  if (p > 0) {
    if (frcs == nullptr) throw std::invalid_argument("forcing functions are NULL");
    if (frc_rows != 1 && frc_rows != p)
      throw std::invalid_argument("forcing functions must have 1 or p rows");
  }
  if (q > 0 && hist_out == nullptr) throw std::invalid_argument("history output is NULL");
}

void finish_timing(bfe_plan_s& H, KernelSampler& sampler, int64_t steps, int64_t launches) {
  BFE_HIP(hipGetLastError());
  BFE_HIP(hipEventRecord(H.ev1, H.stream));
  BFE_HIP(hipEventSynchronize(H.ev1));
  float ms = 0.f;
  BFE_HIP(hipEventElapsedTime(&ms, H.ev0, H.ev1));
  H.stats.loop_ms = ms;
  H.stats.steps = steps;
  H.stats.launches = launches;
  sampler.collect();
}

// Device -> caller's host buffer (FP64), returning the number of non-finite
// values. FP64 results are copied straight into the caller's buffer.
template <typename T>
int64_t copy_out(const DevBuf& d, double* out, size_t count) {
  if constexpr (std::is_same_v<T, double>) {
    BFE_HIP(hipMemcpy(out, d.p, count * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::unique_ptr<T[]> tmp(new T[count]);
    BFE_HIP(hipMemcpy(tmp.get(), d.p, count * sizeof(T), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < count; ++i) out[i] = static_cast<double>(tmp[i]);
  }
  int64_t bad = 0;
  for (size_t i = 0; i < count; ++i) bad += !std::isfinite(out[i]);
  return bad;
}

// Checkpoint segments of the marched range [t0, n_t): counted back from the
// end in steps of C, so every segment but the first has exactly C steps.
struct Segments {
  std::vector<int64_t> start;  // nseg + 1 boundaries, start[nseg] = n_t
  int64_t C = 0;
  int64_t count() const { return static_cast<int64_t>(start.size()) - 1; }
  int64_t len(int64_t k) const { return start[k + 1] - start[k]; }
};

Segments make_segments(int64_t t0, int64_t n_t, int64_t C) {
  Segments S;
  S.C = C;
  const int64_t L = n_t - t0;
  const int64_t nseg = (L + C - 1) / C;
  const int64_t first = L - (nseg - 1) * C;
  S.start.push_back(t0);
  for (int64_t k = 1; k <= nseg; ++k) S.start.push_back(t0 + first + (k - 1) * C);
  return S;
}

// What a checkpointed forward leaves on the handle for bfe_adjoint.
struct AdjointRecord {
  bool valid = false;
  int32_t n_t = 0, p = 0, frc_rows = 0, q = 0;
  int64_t t0 = -1;
  Segments seg;
  std::vector<int32_t> frc_idx, hist_idx;
  std::vector<double> F;  // frc_rows x n_t
  DevBuf ckpt;            // per segment k: u(start_k - 1), u(start_k - 2)
  DevBuf useg[2];         // C x n forward states of one segment each
  int cur = 0;            // useg[cur] holds the last segment after the forward
  // batched record (bfe_run_checkpointed_batched): S sources, states n x S
  // (source axis contiguous); useg[0] holds the last segment's C x n x S
  bool batched = false;
  int32_t S = 1;
  std::vector<int32_t> frc_src;
};

template <typename T>
void launch_step(const bfe_plan_s& H, int S, int64_t nblk, const T* u1, const T* u2, T* u,
                 const BlockLists& L, const T* falpha, const T* F, int32_t F_rows, int32_t t,
                 int32_t q, T* hp, T* fp, int32_t fld_is_vel, T dt, const IfaceRows<T>& IR,
                 T* store0, T* store1) {
  hipStream_t st = H.stream;
  const T* beta = H.damped ? H.beta.as<T>() : nullptr;
  const bool nt = H.nnz_sell * static_cast<int64_t>(sizeof(T) + sizeof(int32_t)) > kStreamBytes;
  auto go1 = [&](auto kern) {
    ColPtrs<T, 1> sp{{store0}};
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(kBlock), 0, st, H.n, nblk, H.off.as<int64_t>(),
                       H.col.as<int32_t>(), H.val.as<T>(), beta, u1, u2, u, L, falpha, F, F_rows,
                       t, q, hp, fp, fld_is_vel, dt, IR, sp);
  };
  auto go2 = [&](auto kern) {
    ColPtrs<T, 2> sp{{store0, store1}};
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(kBlock), 0, st, H.n, nblk, H.off.as<int64_t>(),
                       H.col.as<int32_t>(), H.val.as<T>(), beta, u1, u2, u, L, falpha, F, F_rows,
                       t, q, hp, static_cast<T*>(nullptr), 0, dt, IR, sp);
  };
  if (S == 1) {
    if (H.damped) nt ? go1(k_step<T, 1, true, true>) : go1(k_step<T, 1, true, false>);
    else nt ? go1(k_step<T, 1, false, true>) : go1(k_step<T, 1, false, false>);
  } else {
    if (H.damped) nt ? go2(k_step<T, 2, true, true>) : go2(k_step<T, 2, true, false>);
    else nt ? go2(k_step<T, 2, false, true>) : go2(k_step<T, 2, false, false>);
  }
}

template <typename T, bool D, int NPE, int DPN, int G>
void launch_ebe_generic_g(const bfe_plan_s& H, const EbeOp& E, const EbeGroup& Gr, int64_t nblk,
                          const T* u1, const T* u2, T* u, const BlockLists& L, const T* falpha,
                          const T* F, int32_t F_rows, int32_t t, int32_t q, T* hp, T* fp,
                          int32_t fld_is_vel, T dt, const IfaceRows<T>& IR) {
  hipLaunchKernelGGL((k_step_ebe<T, D, NPE, DPN, G>), dim3(nblk), dim3(kBlock), 0, H.stream,
                     Gr.n_nodes, nblk, Gr.uniform ? nullptr : Gr.dof0.as<int32_t>(), Gr.n_slot,
                     Gr.adj.as<int32_t>(), Gr.nodes.as<int4>(), Gr.scale.as<T>(),
                     Gr.srows.as<T>(), E.alpha_of<T>(Gr), D ? E.beta_of<T>(Gr) : nullptr, u1, u2,
                     u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR, Gr.coef_node);
}

template <typename T, bool D, int NPE, int DPN>
void launch_ebe_generic(const bfe_plan_s& H, const EbeOp& E, const EbeGroup& Gr, int64_t nblk,
                        const T* u1, const T* u2, T* u, const BlockLists& L, const T* falpha,
                        const T* F, int32_t F_rows, int32_t t, int32_t q, T* hp, T* fp,
                        int32_t fld_is_vel, T dt, const IfaceRows<T>& IR) {
  switch (H.ebe_group) {
    case 2: launch_ebe_generic_g<T, D, NPE, DPN, 2>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); break;
    case 4: launch_ebe_generic_g<T, D, NPE, DPN, 4>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); break;
    default: launch_ebe_generic_g<T, D, NPE, DPN, 8>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); break;
  }
}

template <typename T, bool D, typename PT, int DPN>
void launch_ebe_pat(const bfe_plan_s& H, const EbeOp& E, const EbeGroup& Gr, int64_t nblk,
                    const T* u1, const T* u2, T* u, const BlockLists& L, const T* falpha,
                    const T* F, int32_t F_rows, int32_t t, int32_t q, T* hp, T* fp,
                    int32_t fld_is_vel, T dt, const IfaceRows<T>& IR) {
  ClassOffsets<PT::NC> co;
  for (int j = 0; j < PT::NC; ++j) co.off[j] = static_cast<int32_t>(Gr.doff[j]);
  AnchorSlots<PT::NS> an{};
  const bool anchored = Gr.anchor.bytes > 0;  // only on groups whose slots are mostly scaled
  if (anchored) {
    for (int k = 0; k < PT::NS; ++k)
      an.base[k] = static_cast<int32_t>(static_cast<int64_t>(Gr.anc_row[k]) * Gr.n_nodes +
                                        Gr.anc_noff[k]);
    an.len = static_cast<int32_t>(Gr.anchor.bytes / sizeof(T));
  }
  const uint32_t* sm = Gr.smask.as<uint32_t>();
  if constexpr (PT::NS <= kAnchorMaxSlots) {
    if (anchored) {
      hipLaunchKernelGGL((k_step_ebe_pat<T, D, PT, DPN, true>), dim3(nblk), dim3(kBlock), 0,
                       H.stream, Gr.n_nodes, nblk, H.n, Gr.uniform ? nullptr : Gr.dof0.as<int32_t>(),
                       co, an, Gr.pmask.as<uint32_t>(), sm, Gr.nbr.as<int32_t>(),
                       Gr.adj.as<int32_t>(), Gr.scale.as<T>(), Gr.anchor.as<T>(),
                       Gr.srows_c.as<T>(), Gr.stencil.as<T>(), E.alpha_of<T>(Gr),
                       D ? E.beta_of<T>(Gr) : nullptr, u1, u2, u, L, falpha, F, F_rows, t, q, hp,
                       fp, fld_is_vel, dt, IR, Gr.coef_node);
      return;
    }
  }
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(kBlock), 0, H.stream, Gr.n_nodes, nblk, H.n,
                       Gr.uniform ? nullptr : Gr.dof0.as<int32_t>(), co, an,
                       Gr.pmask.as<uint32_t>(), sm, Gr.nbr.as<int32_t>(), Gr.adj.as<int32_t>(),
                       Gr.scale.as<T>(), static_cast<const T*>(nullptr), Gr.srows_c.as<T>(),
                       Gr.stencil.as<T>(), E.alpha_of<T>(Gr), D ? E.beta_of<T>(Gr) : nullptr, u1,
                       u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR,
                       Gr.coef_node);
  };
  // structurally sparse slot rows (checked at set-up): zero products skipped
  if (std::is_same<PT, PatKuhn24>::value && Gr.sparse)
    go(k_step_ebe_pat<T, D, PT, DPN, false, true>);
  else
    go(k_step_ebe_pat<T, D, PT, DPN, false, false>);
}

// Whether a kernel exists for the group's shape (generic) or pattern.
bool pattern_has_anchor(int pattern) {
  return (pattern == PatTriRect::ID && PatTriRect::NS <= kAnchorMaxSlots) ||
         (pattern == PatTriIso::ID && PatTriIso::NS <= kAnchorMaxSlots);
}
bool has_generic_kernel(int npe, int dpn) {
  return (npe == 3 && dpn >= 1 && dpn <= 3) || (npe == 4 && dpn == 3);
}
bool has_pattern_kernel(int pattern, int dpn) {
  if (pattern == PatKuhn24::ID) return dpn == 3;
  return (pattern == PatTriRect::ID || pattern == PatTriIso::ID) && dpn >= 1 && dpn <= 3;
}

// One group's step kernel (pattern kernel unless generic or switched off).
template <typename T, bool D>
void launch_ebe_one(const bfe_plan_s& H, const EbeOp& E, const EbeGroup& Gr, int64_t nblk,
                    const T* u1, const T* u2, T* u, const BlockLists& L, const T* falpha,
                    const T* F, int32_t F_rows, int32_t t, int32_t q, T* hp, T* fp,
                    int32_t fld_is_vel, T dt, const IfaceRows<T>& IR, bool generic) {
  if (!generic && H.ebe_kernel && Gr.pattern != 0) {
    switch (Gr.pattern * 10 + Gr.dpn) {
      case PatKuhn24::ID * 10 + 3: launch_ebe_pat<T, D, PatKuhn24, 3>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
      case PatTriRect::ID * 10 + 1: launch_ebe_pat<T, D, PatTriRect, 1>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
      case PatTriRect::ID * 10 + 2: launch_ebe_pat<T, D, PatTriRect, 2>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
      case PatTriRect::ID * 10 + 3: launch_ebe_pat<T, D, PatTriRect, 3>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
      case PatTriIso::ID * 10 + 1: launch_ebe_pat<T, D, PatTriIso, 1>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
      case PatTriIso::ID * 10 + 2: launch_ebe_pat<T, D, PatTriIso, 2>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
      case PatTriIso::ID * 10 + 3: launch_ebe_pat<T, D, PatTriIso, 3>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
      default: throw std::invalid_argument("no pattern kernel for this element group");
    }
  }
  switch (Gr.npe * 10 + Gr.dpn) {
    case 31: launch_ebe_generic<T, D, 3, 1>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
    case 32: launch_ebe_generic<T, D, 3, 2>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
    case 33: launch_ebe_generic<T, D, 3, 3>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
    case 43: launch_ebe_generic<T, D, 4, 3>(H, E, Gr, nblk, u1, u2, u, L, falpha, F, F_rows, t, q, hp, fp, fld_is_vel, dt, IR); return;
    default: throw std::invalid_argument("unsupported element group shape");
  }
}

// Per-group lists of one run (group-local rows).
struct EbeRunLists {
  DevBuf fptr, frow, fsrc, ffrow, falpha, mptr, mrow, mk;
};

// Grid of one group: one thread per node.
inline int64_t ebe_blocks(const EbeGroup& Gr) { return (Gr.n_nodes + kBlock - 1) / kBlock; }

// Builds the forcing and monitor lists of every group for one run (the same
// last-wins deduplication as make_run_buffers).
template <typename T>
std::vector<EbeRunLists> ebe_run_lists(const bfe_plan_s& H, const EbeOp& E, const int32_t* frc_idx,
                                       int32_t p, int32_t frc_rows, const int32_t* hist_idx,
                                       int32_t q) {
  std::vector<std::pair<int32_t, int32_t>> key;  // (dof, entry)
  for (int32_t i = 0; i < p; ++i) key.push_back({frc_idx[i], i});
  std::stable_sort(key.begin(), key.end(), [](auto& a, auto& b) { return a.first < b.first; });
  const size_t ng = E.groups.size();
  std::vector<std::vector<int32_t>> frows(ng), fsrc(ng), ffrow(ng), mrows(ng), mks(ng);
  std::vector<std::vector<double>> falpha(ng);
  for (size_t i = 0; i < key.size(); ++i)
    if (i + 1 == key.size() || key[i + 1].first != key[i].first) {
      const int32_t r = key[i].first;
      const int g = E.row_group[r];
      if (g < 0) throw std::invalid_argument("forcing on a DOF that no element group owns");
      frows[g].push_back(E.row_local[r]);
      fsrc[g].push_back(0);
      ffrow[g].push_back(frc_rows == 1 ? 0 : key[i].second);
      falpha[g].push_back(H.alpha[r]);
    }
  for (int32_t k = 0; k < q; ++k) {
    const int32_t r = hist_idx[k];
    const int g = E.row_group[r];
    if (g < 0) throw std::invalid_argument("monitor on a DOF that no element group owns");
    mrows[g].push_back(E.row_local[r]);
    mks[g].push_back(k);
  }
  std::vector<EbeRunLists> out(ng);
  for (size_t g = 0; g < ng; ++g) {
    const EbeGroup& Gr = E.groups[g];
    const int64_t nb = ebe_blocks(Gr), rpb = static_cast<int64_t>(kBlock) * Gr.dpn;
    Grouped gf = group_by_block(nb, rpb, frows[g], &fsrc[g], &ffrow[g]);
    // falpha follows the grouped order: regroup it with the rows as payload index
    std::vector<int32_t> idx(frows[g].size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = static_cast<int32_t>(i);
    Grouped gfi = group_by_block(nb, rpb, frows[g], &idx, nullptr);
    std::vector<double> fa(gfi.a.size());
    for (size_t i = 0; i < fa.size(); ++i) fa[i] = falpha[g][gfi.a[i]];
    Grouped gm = group_by_block(nb, rpb, mrows[g], &mks[g], nullptr);
    EbeRunLists& R = out[g];
    R.fptr = upload(gf.ptr); R.frow = upload(gf.row); R.fsrc = upload(gf.a);
    R.ffrow = upload(gf.b); R.falpha = upload(cast_vec<T>(fa));
    R.mptr = upload(gm.ptr); R.mrow = upload(gm.row); R.mk = upload(gm.a);
  }
  return out;
}

inline BlockLists ebe_lists(const EbeRunLists& R, const EbeGroup& Gr) {
  BlockLists L;
  L.fptr = R.fptr.as<int32_t>(); L.f_row = R.frow.as<int32_t>(); L.f_src = R.fsrc.as<int32_t>();
  L.f_frow = R.ffrow.as<int32_t>();
  L.mptr = R.mptr.as<int32_t>(); L.m_row = R.mrow.as<int32_t>(); L.m_k = R.mk.as<int32_t>();
  L.iptr = Gr.iptr.as<int32_t>(); L.i_row = Gr.irow.as<int32_t>(); L.i_k = Gr.ik.as<int32_t>();
  return L;
}

// One step of every group (each group writes only its own rows).
template <typename T>
void launch_step_ebe(const bfe_plan_s& H, const EbeOp& E, const std::vector<EbeRunLists>* RL,
                     const BlockLists* L0, const T* u1, const T* u2, T* u, const T* F,
                     int32_t F_rows, int32_t t, int32_t q, T* hp, T* fp, int32_t fld_is_vel, T dt,
                     const IfaceRows<T>& IR, bool generic = false) {
  for (size_t g = 0; g < E.groups.size(); ++g) {
    const EbeGroup& Gr = E.groups[g];
    const BlockLists L = RL ? ebe_lists((*RL)[g], Gr) : *L0;
    const T* falpha = RL ? (*RL)[g].falpha.as<T>() : nullptr;
    if (H.damped)
      launch_ebe_one<T, true>(H, E, Gr, ebe_blocks(Gr), u1, u2, u, L, falpha, F, F_rows, t, q, hp,
                              fp, fld_is_vel, dt, IR, generic);
    else
      launch_ebe_one<T, false>(H, E, Gr, ebe_blocks(Gr), u1, u2, u, L, falpha, F, F_rows, t, q, hp,
                               fp, fld_is_vel, dt, IR, generic);
  }
}

// One step of both operators from the same random state (u2 = 0, no
// forcing, monitors or interface correction): returns
// max |u_ebe - u_csr| / max |u_csr| over the rows the groups own.
template <typename T>
double compare_ebe_with_csr(const bfe_plan_s& H, const EbeOp& E, bool generic) {
  const int64_t n = H.n;
  const int64_t nb_c = (n + kRowsSingle - 1) / kRowsSingle;
  int64_t nb_max = nb_c;
  for (const auto& g : E.groups) nb_max = std::max(nb_max, ebe_blocks(g));
  std::vector<int32_t> zeros(nb_max + 1, 0);
  DevBuf z = upload(zeros);
  BlockLists L0;
  L0.fptr = L0.f_row = L0.f_src = L0.f_frow = z.as<int32_t>();
  L0.mptr = L0.m_row = L0.m_k = z.as<int32_t>();
  L0.iptr = L0.i_row = L0.i_k = z.as<int32_t>();
  std::vector<T> x(n);
  uint64_t st = 0x9E3779B97F4A7C15ull;
  for (int64_t i = 0; i < n; ++i) {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    x[i] = static_cast<T>(static_cast<double>(st >> 11) * 0x1.0p-53 * 2.0 - 1.0);
  }
  DevBuf dx = upload(x);
  DevBuf buf(sizeof(T) * static_cast<size_t>(n) * 3);
  BFE_HIP(hipMemsetAsync(buf.p, 0, buf.bytes, H.stream));
  T* u2 = buf.as<T>();
  T* y1 = u2 + n;
  T* y2 = u2 + 2 * n;
  IfaceRows<T> I{};
  I.wI = nullptr;
  const T dt = static_cast<T>(H.dt);
  launch_step<T>(H, 1, nb_c, dx.as<T>(), u2, y1, L0, nullptr, nullptr, 1, 0, 0, nullptr, nullptr,
                 0, dt, I, nullptr, nullptr);
  launch_step_ebe<T>(H, E, nullptr, &L0, dx.as<T>(), u2, y2, nullptr, 1, 0, 0, nullptr, nullptr,
                     0, dt, I, generic);
  BFE_HIP(hipGetLastError());
  BFE_HIP(hipStreamSynchronize(H.stream));
  std::vector<T> a(n), b(n);
  BFE_HIP(hipMemcpy(a.data(), y1, sizeof(T) * n, hipMemcpyDeviceToHost));
  BFE_HIP(hipMemcpy(b.data(), y2, sizeof(T) * n, hipMemcpyDeviceToHost));
  double dmax = 0.0, amax = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double d = std::fabs(static_cast<double>(a[i]) - static_cast<double>(b[i]));
    if (!(d <= dmax)) dmax = d;  // NaN propagates
    amax = std::max(amax, std::fabs(static_cast<double>(a[i])));
  }
  return amax > 0.0 ? dmax / amax : dmax;
}

// Owns a stream capture and its graph: a capture still open when an error
// unwinds is ended, so the stream is usable again.
struct GraphCapture {
  hipStream_t stream = nullptr;  // set while capturing
  hipStream_t launched = nullptr;  // stream the instantiated graph was launched on
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  GraphCapture() = default;
  GraphCapture(const GraphCapture&) = delete;
  GraphCapture& operator=(const GraphCapture&) = delete;
  ~GraphCapture() {
    if (stream) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
    }
    if (exec) {
      if (launched) (void)hipStreamSynchronize(launched);
      (void)hipGraphExecDestroy(exec);
    }
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// Element and node tables of the per-element field output.
struct FieldElements {
  int64_t n_el = 0, n_nodes = 0;
  int32_t npe = 0, max_df = 0;
  DevBuf els, node_dof, m;
  int32_t dim = 0;  // node coordinates (BFE_FIELD_DIV / _CURL), 0: not set
  DevBuf nds;
};

template <typename T>
void run_single(bfe_plan_s& H, int32_t n_t, const int32_t* frc_idx, int32_t p,
                const double* frcs, int32_t frc_rows, const int32_t* hist_idx, int32_t q,
                double* hist_out, int32_t field_every, int32_t field_is_vel,
                double* field_out, int32_t* ti_start_out, AdjointRecord* rec, int32_t ckpt_every,
                double* el_ke_out = nullptr, int32_t field_type = BFE_FIELD_KE,
                int32_t field_comp = 0) {
  const int64_t n = H.n;
  validate_run(n, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out);
  const int32_t n_frames = field_every > 0 ? (n_t - 1) / field_every + 1 : 0;
  auto FE = std::static_pointer_cast<FieldElements>(H.field_elems);
  if (el_ke_out != nullptr && !FE) throw std::invalid_argument("field elements not set");
  if (n_frames > 0 && field_out == nullptr && el_ke_out == nullptr)
    throw std::invalid_argument("field output is NULL");

  H.stats = bfe_run_stats{};
  H.stats.columns = 1;
  const int64_t t0 = p > 0 ? first_forced_step(frcs, frc_rows, n_t) : -1;
  if (ti_start_out) *ti_start_out = t0 < 0 ? n_t + 1 : static_cast<int32_t>(t0 + 1);
  if (rec) {
    *rec = AdjointRecord{};
    rec->n_t = n_t;
    rec->p = p;
    rec->frc_rows = frc_rows;
    rec->q = q;
    rec->t0 = t0;
    rec->frc_idx.assign(frc_idx, frc_idx + p);
    rec->hist_idx.assign(hist_idx, hist_idx + q);
    if (p > 0) rec->F.assign(frcs, frcs + static_cast<size_t>(frc_rows) * n_t);
    rec->valid = true;
  }
  if (t0 < 0) {  // empty loop (v6:163-180): outputs are zero
    if (q > 0) std::memset(hist_out, 0, sizeof(double) * static_cast<size_t>(q) * n_t);
    if (n_frames > 0 && field_out != nullptr)
      std::memset(field_out, 0, sizeof(double) * static_cast<size_t>(n) * n_frames);
    if (n_frames > 0 && el_ke_out != nullptr)
      std::memset(el_ke_out, 0, sizeof(double) * static_cast<size_t>(FE->n_el) * n_frames);
    return;
  }

  BFE_HIP(hipSetDevice(H.device));
  hipStream_t st = H.stream;
  // the checkpointed forward stays on the CSR operator that the adjoint's
  // recomputation uses
  const EbeOp* E = (H.ebe && H.use_ebe && rec == nullptr) ? H.ebe.get() : nullptr;
  const int64_t nblk = (n + kRowsSingle - 1) / kRowsSingle;
  RunBuffers R;
  make_run_buffers<T>(R, H, kRowsSingle, nblk, 1, frc_idx, nullptr, p, frcs, frc_rows, n_t,
                      hist_idx, q);
  const BlockLists L = R.lists(H);
  std::vector<EbeRunLists> RL;  // per element group, group-local rows
  if (E) RL = ebe_run_lists<T>(H, *E, frc_idx, p, frc_rows, hist_idx, q);
  DevBuf dhist(sizeof(T) * static_cast<size_t>(q) * n_t);
  DevBuf dfld(sizeof(T) * static_cast<size_t>(n) * n_frames);
  DevBuf ubuf(sizeof(T) * static_cast<size_t>(n) * 3);
  BFE_HIP(hipMemsetAsync(ubuf.p, 0, ubuf.bytes, st));
  if (dhist.bytes) BFE_HIP(hipMemsetAsync(dhist.p, 0, dhist.bytes, st));
  if (dfld.bytes) BFE_HIP(hipMemsetAsync(dfld.p, 0, dfld.bytes, st));
  T *u = ubuf.as<T>(), *u1 = u + n, *u2 = u + 2 * n;
  T* hist = dhist.as<T>();
  T* fld = dfld.as<T>();
  T* wI = R.wI.as<T>();
  const IfaceRows<T> IR = iface_rows<T>(H, wI);
  const T* falpha = R.falpha.as<T>();
  const T* F = R.F.as<T>();
  const T dtT = static_cast<T>(H.dt);
  int64_t next_seg = 0;
  if (rec) {
    rec->seg = make_segments(t0, n_t, ckpt_every);
    const int64_t nseg = rec->seg.count();
    rec->ckpt = DevBuf(sizeof(T) * static_cast<size_t>(n) * 2 * nseg);
    for (auto& b : rec->useg) b = DevBuf(sizeof(T) * static_cast<size_t>(n) * rec->seg.C);
    rec->cur = 0;
  }
  int64_t launches = 0;
  // persistent mode: one cooperative launch runs the loop (CSR operator,
  // plain marches); graph mode: the whole time loop is captured into one
  // hipGraph and then launched (plain marches only). Neither samples
  // per-step kernels.
  const bool persistent = H.use_persistent != 0 && E == nullptr && rec == nullptr;
  const bool graph = H.use_graph != 0 && rec == nullptr && !persistent;
  KernelSampler sampler(H, t0, (graph || persistent) ? 0 : n_t - t0);
  GraphCapture cap;
  if (graph) {
    BFE_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    cap.stream = st;
  } else {
    BFE_HIP(hipEventRecord(H.ev0, st));
  }
  if (persistent) {
    launch_persistent<T>(H, nblk, t0, n_t, ubuf.as<T>(), L, falpha, F, frc_rows, q, hist,
                         n_frames > 0 ? field_every : 0, fld, field_is_vel, dtT, IR);
    ++launches;
    const int64_t k = (n_t - t0) % 3;  // the rotation after the last step
    u = ubuf.as<T>() + ((3 - k) % 3) * n;
    u1 = ubuf.as<T>() + ((4 - k) % 3) * n;
    u2 = ubuf.as<T>() + ((5 - k) % 3) * n;
  }
  for (int64_t t = persistent ? n_t : t0; t < n_t; ++t) {
    const int64_t prev = t - 1;
    T* hp = (q > 0 && prev >= t0) ? hist + prev * q : nullptr;
    T* fp = (n_frames > 0 && prev >= t0 && prev % field_every == 0)
                ? fld + (prev / field_every) * n : nullptr;
    T* store = nullptr;
    if (rec) {
      const Segments& sg = rec->seg;
      if (next_seg < sg.count() && t == sg.start[next_seg]) {
        // checkpoint: the two states the step at t starts from
        T* ck = rec->ckpt.as<T>() + next_seg * 2 * n;
        BFE_HIP(hipMemcpyAsync(ck, u1, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
        BFE_HIP(hipMemcpyAsync(ck + n, u2, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
        ++next_seg;
      }
      const int64_t last = sg.count() - 1;
      if (t >= sg.start[last]) store = rec->useg[0].as<T>() + (t - sg.start[last]) * n;
    }
    if (!graph) sampler.before(t);
    if (E)
      launch_step_ebe<T>(H, *E, &RL, nullptr, u1, u2, u, F, frc_rows, static_cast<int32_t>(t), q,
                         hp, fp, field_is_vel, dtT, IR);
    else
      launch_step<T>(H, 1, nblk, u1, u2, u, L, falpha, F, frc_rows, static_cast<int32_t>(t), q,
                     hp, fp, field_is_vel, dtT, IR, store, nullptr);
    if (!graph) sampler.after(t);
    ++launches;
    if (H.nI > 0 && H.has_xinv) {
      launch_iface<T>(H, 1, wI, u);
      ++launches;
      if (store) BFE_HIP(hipMemcpyAsync(store, u, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    }
    T* tmp = u2; u2 = u1; u1 = u; u = tmp;
  }
  {
    const int64_t last = n_t - 1;
    T* hp = q > 0 ? hist + last * q : nullptr;
    T* fp = (n_frames > 0 && last % field_every == 0) ? fld + (last / field_every) * n : nullptr;
    const int64_t work = std::max<int64_t>(fp ? n : 0, q);
    if (work > 0) {
      hipLaunchKernelGGL((k_snap<T>), dim3((work + 255) / 256), dim3(256), 0, st, n, 1, u1, u2,
                         q, R.hidx.as<int32_t>(), hp, fp, field_is_vel, dtT);
      ++launches;
    }
  }
  if (graph) {
    BFE_HIP(hipGetLastError());
    cap.stream = nullptr;
    BFE_HIP(hipStreamEndCapture(st, &cap.graph));
    BFE_HIP(hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0));
    BFE_HIP(hipEventRecord(H.ev0, st));
    cap.launched = st;
    BFE_HIP(hipGraphLaunch(cap.exec, st));
  }
  finish_timing(H, sampler, n_t - t0, launches);
  if (q > 0) H.stats.nonfinite = copy_out<T>(dhist, hist_out, static_cast<size_t>(q) * n_t);
  if (n_frames > 0 && field_out != nullptr)
    copy_out<T>(dfld, field_out, static_cast<size_t>(n) * n_frames);
  if (n_frames > 0 && el_ke_out != nullptr) {
    // per-element field on the device: only n_el x F values come back
    const int64_t work = FE->n_el * n_frames;
    DevBuf dke(sizeof(T) * static_cast<size_t>(work));
    if (field_type == BFE_FIELD_KE)
      hipLaunchKernelGGL((k_field_ke<T>), dim3((work + 255) / 256), dim3(256), 0, st, FE->n_el,
                         FE->npe, FE->els.as<int32_t>(), FE->max_df, FE->node_dof.as<int32_t>(),
                         FE->m.as<T>(), fld, n, n_frames, dke.as<T>());
    else
      hipLaunchKernelGGL((k_field_reduce<T>), dim3((work + 255) / 256), dim3(256), 0, st,
                         field_type, field_comp, FE->n_el, FE->npe, FE->els.as<int32_t>(),
                         FE->max_df, FE->node_dof.as<int32_t>(), FE->nds.as<double>(), FE->dim,
                         fld, n, n_frames, dke.as<T>());
    BFE_HIP(hipGetLastError());
    BFE_HIP(hipStreamSynchronize(st));
    copy_out<T>(dke, el_ke_out, static_cast<size_t>(work));
  }
}

// Element data for the parameter gradients.
struct GradElements {
  int32_t n_el = 0, kmax = 0;
  DevBuf dofs, K, m, w;
  // node-pair form of the batched gradient (pair_gd = 0: not available)
  int32_t pair_gd = 0, pair_em = 0, pair_kv = 0, pair_stride = 0;
  int64_t n_groups = 0;
  DevBuf gdof, gnb, egrp, eoff;  // node first DOF, upper neighbours, element nodes and blocks
};

// Node-pair accumulation over one segment (k_pair_acc) and the final
// contraction (k_pair_contract), for the shapes that have kernels.
bool pair_shape_ok(int gd, int em, int kv) {
  const bool acc_ok = (gd == 1 && em >= 2 && em <= 6) || (gd == 2 && em >= 2 && em <= 4);
  const bool con_ok = (gd == 1 && (kv == 2 || kv == 3 || kv == 4)) || (gd == 2 && (kv == 2 || kv == 3 || kv == 4));
  return acc_ok && con_ok;
}

// Node-pair layout of gradient elements (see k_pair_acc): the vertices of
// every element are GD consecutive DOFs starting at a node's first DOF (or
// all absent), each DOF belongs to one node, and no node has more than
// em_max neighbours above it. Returns false when the elements do not have
// that form; the batched adjoint then uses k_grad_multi.
bool pair_layout(const std::vector<int32_t>& d, int64_t ne, int32_t kmax, int64_t n, int gd,
                 int em_max, GradElements& G) {
  if (kmax % gd != 0) return false;
  const int kv = kmax / gd;
  std::vector<int32_t> start(n, -1);  // DOF -> first DOF of its node
  std::vector<int8_t> comp(n, -1);
  for (int64_t e = 0; e < ne; ++e)
    for (int v = 0; v < kv; ++v) {
      const int32_t* x = &d[e * kmax + v * gd];
      if (x[0] < 0) {
        for (int c = 1; c < gd; ++c)
          if (x[c] >= 0) return false;
        continue;
      }
      for (int c = 0; c < gd; ++c) {
        if (x[c] != x[0] + c) return false;
        if (start[x[c]] < 0) {
          start[x[c]] = x[0];
          comp[x[c]] = static_cast<int8_t>(c);
        } else if (start[x[c]] != x[0] || comp[x[c]] != c) {
          return false;
        }
      }
    }
  std::vector<int32_t> gid(n, -1), gdof;
  for (int64_t i = 0; i < n; ++i)
    if (start[i] == i) {
      gid[i] = static_cast<int32_t>(gdof.size());
      gdof.push_back(static_cast<int32_t>(i));
    }
  const int64_t ng = static_cast<int64_t>(gdof.size());
  if (ng == 0) return false;
  std::vector<int32_t> egrp(static_cast<size_t>(ne) * kv, -1);
  std::vector<uint64_t> pairs;
  for (int64_t e = 0; e < ne; ++e) {
    for (int v = 0; v < kv; ++v) {
      const int32_t x = d[e * kmax + v * gd];
      egrp[e * kv + v] = x >= 0 ? gid[x] : -1;
    }
    for (int i = 0; i < kv; ++i)
      for (int j = i + 1; j < kv; ++j) {
        const int32_t A = egrp[e * kv + i], B = egrp[e * kv + j];
        if (A < 0 || B < 0 || A == B) continue;
        const uint64_t lo = static_cast<uint64_t>(std::min(A, B)), hi = static_cast<uint64_t>(std::max(A, B));
        pairs.push_back((lo << 32) | hi);
      }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  std::vector<int32_t> deg(ng, 0);
  for (uint64_t p : pairs) ++deg[p >> 32];
  int em = 1;
  for (int32_t x : deg) em = std::max(em, static_cast<int>(x));
  if (em > em_max) return false;
  em = std::max(em, 2);
  if (!pair_shape_ok(gd, em, kv)) return false;
  // upper neighbours, ascending, slot-major; slot of each pair
  std::vector<int32_t> gnb(static_cast<size_t>(em) * ng, -1), fill(ng, 0);
  for (uint64_t p : pairs) {
    const int32_t A = static_cast<int32_t>(p >> 32), B = static_cast<int32_t>(p & 0xffffffffu);
    gnb[static_cast<int64_t>(fill[A]++) * ng + A] = B;
  }
  auto slot_of = [&](int32_t A, int32_t B) {
    for (int k = 0; k < em; ++k)
      if (gnb[static_cast<int64_t>(k) * ng + A] == B) return k;
    return -1;
  };
  const int bs = gd * gd;
  const int stride = bs + 2 * gd + em * 2 * bs;
  std::vector<int64_t> eoff(static_cast<size_t>(ne) * kv * kv, -1);
  for (int64_t e = 0; e < ne; ++e)
    for (int i = 0; i < kv; ++i)
      for (int j = 0; j < kv; ++j) {
        const int32_t A = egrp[e * kv + i], B = egrp[e * kv + j];
        if (A < 0 || B < 0) continue;
        int64_t off;
        if (A == B) off = static_cast<int64_t>(A) * stride;
        else if (A < B) off = static_cast<int64_t>(A) * stride + bs + 2 * gd + slot_of(A, B) * 2 * bs;
        else off = static_cast<int64_t>(B) * stride + bs + 2 * gd + slot_of(B, A) * 2 * bs + bs;
        eoff[(e * kv + i) * kv + j] = off;
      }
  G.pair_gd = gd;
  G.pair_em = em;
  G.pair_kv = kv;
  G.pair_stride = stride;
  G.n_groups = ng;
  G.gdof = upload(gdof);
  G.gnb = upload(gnb);
  G.egrp = upload(egrp);
  G.eoff = upload(eoff);
  return true;
}

template <typename T>
void launch_pair_acc(const bfe_plan_s& H, const GradElements& G, int32_t S, const T* U,
                     int64_t us, int64_t ud, const T* Lam, int64_t ls, int64_t ld, int32_t Ls,
                     const T* um1, const T* um2, double* acc) {
  int32_t GS = 1;
  while (GS < S) GS <<= 1;
  const int64_t threads = G.n_groups * GS;
  const dim3 grid((threads + kBlock - 1) / kBlock), blk(kBlock);
  const int32_t *gd = G.gdof.as<int32_t>(), *nb = G.gnb.as<int32_t>();
#define BFE_PAIR(GD_, EM_)                                                                     case GD_ * 16 + EM_:                                                                           hipLaunchKernelGGL((k_pair_acc<T, GD_, EM_>), grid, blk, 0, H.stream, G.n_groups, S, GS,                        gd, nb, U, us, ud, Lam, ls, ld, Ls, um1, um2, acc);                       break;
  switch (G.pair_gd * 16 + G.pair_em) {
    BFE_PAIR(1, 2) BFE_PAIR(1, 3) BFE_PAIR(1, 4) BFE_PAIR(1, 5) BFE_PAIR(1, 6)
    BFE_PAIR(2, 2) BFE_PAIR(2, 3) BFE_PAIR(2, 4)
    default: throw std::invalid_argument("no node-pair kernel for this element layout");
  }
#undef BFE_PAIR
}

template <typename T>
void launch_pair_contract(const bfe_plan_s& H, const GradElements& G, const double* acc,
                          double* gK, double* gM) {
  const dim3 grid((G.n_el + kBlock - 1) / kBlock), blk(kBlock);
  const double dt2 = H.dt * H.dt;
  const T *Kp = G.K.as<T>(), *mp = G.m.as<T>(), *wp = G.w.as<T>();
#define BFE_PCON(KV_, GD_)                                                                       case KV_ * 16 + GD_:                                                                             hipLaunchKernelGGL((k_pair_contract<T, KV_, GD_>), grid, blk, 0, H.stream, G.n_el,                                G.pair_stride, G.egrp.as<int32_t>(), G.eoff.as<int64_t>(), Kp, mp, wp,                        acc, dt2, gK, gM);                                                          break;
  switch (G.pair_kv * 16 + G.pair_gd) {
    BFE_PCON(2, 1) BFE_PCON(3, 1) BFE_PCON(4, 1) BFE_PCON(2, 2) BFE_PCON(3, 2) BFE_PCON(4, 2)
    default: throw std::invalid_argument("no node-pair contraction for this element layout");
  }
#undef BFE_PCON
}

template <typename T>
void launch_grad(const bfe_plan_s& H, const GradElements& G, const T* U, const T* Lam, int32_t Ls,
                 const T* um1, const T* um2, double* gK, double* gM) {
  const dim3 grid((G.n_el + kBlock - 1) / kBlock), blk(kBlock);
  const double dt2 = H.dt * H.dt;
  const int32_t* d = G.dofs.as<int32_t>();
  const T *Kp = G.K.as<T>(), *mp = G.m.as<T>(), *wp = G.w.as<T>();
#define BFE_GRAD(KM)                                                                       \
  hipLaunchKernelGGL((k_grad<T, KM>), grid, blk, 0, H.stream, G.n_el, d, Kp, mp, wp, H.n, U, \
                     Lam, Ls, um1, um2, dt2, gK, gM)
  switch (G.kmax) {
    case 2: BFE_GRAD(2); break;
    case 3: BFE_GRAD(3); break;
    case 4: BFE_GRAD(4); break;
    case 6: BFE_GRAD(6); break;
    case 8: BFE_GRAD(8); break;
    case 9: BFE_GRAD(9); break;
    case 12: BFE_GRAD(12); break;
    default: throw std::invalid_argument("unsupported element DOF count for gradients");
  }
#undef BFE_GRAD
}

// Reverse sweep: for k = last .. 0, the adjoint over segment k runs in
// column 1 of a two-column march while column 0 recomputes segment k-1 from
// its checkpoint; then the gradient kernel consumes segment k.
template <typename T>
void launch_grad_multi(const bfe_plan_s& H, const GradElements& G, int32_t S, const T* U,
                       int64_t us, int64_t ud, const T* Lam, int64_t ls, int64_t ld, int32_t Ls,
                       const T* um1, const T* um2, double* gK, double* gM) {
  int32_t GS = 1;
  while (GS < S) GS <<= 1;
  const int64_t threads = static_cast<int64_t>(G.n_el) * GS;
  const dim3 grid((threads + kBlock - 1) / kBlock), blk(kBlock);
  const double dt2 = H.dt * H.dt;
  const int32_t* d = G.dofs.as<int32_t>();
  const T *Kp = G.K.as<T>(), *mp = G.m.as<T>(), *wp = G.w.as<T>();
#define BFE_GRADM(KM)                                                                          \
  hipLaunchKernelGGL((k_grad_multi<T, KM>), grid, blk, 0, H.stream, G.n_el, S, GS, d, Kp, mp, wp, \
                     U, us, ud, Lam, ls, ld, Ls, um1, um2, dt2, gK, gM)
  switch (G.kmax) {
    case 2: BFE_GRADM(2); break;
    case 3: BFE_GRADM(3); break;
    case 4: BFE_GRADM(4); break;
    case 6: BFE_GRADM(6); break;
    case 8: BFE_GRADM(8); break;
    case 9: BFE_GRADM(9); break;
    case 12: BFE_GRADM(12); break;
    default: throw std::invalid_argument("unsupported element DOF count for gradients");
  }
#undef BFE_GRADM
}

template <typename T>
void run_adjoint(bfe_plan_s& H, AdjointRecord& rec, const GradElements& G,
                 const double* hist_grad, double* g_stiff, double* g_dens) {
  const int64_t n = H.n;
  std::memset(g_stiff, 0, sizeof(double) * G.n_el);
  std::memset(g_dens, 0, sizeof(double) * G.n_el);
  H.stats = bfe_run_stats{};
  H.stats.columns = 2;
  if (rec.t0 < 0) return;
  BFE_HIP(hipSetDevice(H.device));
  hipStream_t st = H.stream;
  const int32_t n_t = rec.n_t;
  const Segments& sg = rec.seg;
  const int64_t nseg = sg.count(), C = sg.C;
  const double dt2 = H.dt * H.dt;

  // adjoint forcing: receivers with duplicate DOFs summed, scaled by 1/dt^2
  std::vector<int32_t> rdof;
  std::vector<int32_t> rrow(rec.q);
  {
    std::vector<std::pair<int32_t, int32_t>> kk;
    for (int32_t k = 0; k < rec.q; ++k) kk.push_back({rec.hist_idx[k], k});
    std::sort(kk.begin(), kk.end());
    for (auto& e : kk) {
      if (rdof.empty() || rdof.back() != e.first) rdof.push_back(e.first);
      rrow[e.second] = static_cast<int32_t>(rdof.size()) - 1;
    }
  }
  const int32_t qa = static_cast<int32_t>(rdof.size());
  std::vector<double> Fadj(static_cast<size_t>(qa) * n_t, 0.0);
  for (int64_t t = 0; t < n_t; ++t)
    for (int32_t k = 0; k < rec.q; ++k)
      Fadj[t * qa + rrow[k]] += hist_grad[t * rec.q + k] / dt2;

  // forcing entries of the two-column march: forward (col 0) then adjoint (col 1)
  const int32_t pf = rec.p, frow_f = rec.frc_rows;
  const int32_t rows_dual = (pf > 0 ? frow_f : 0) + qa;
  std::vector<int32_t> fi, fs;
  for (int32_t i = 0; i < pf; ++i) { fi.push_back(rec.frc_idx[i]); fs.push_back(0); }
  for (int32_t k = 0; k < qa; ++k) { fi.push_back(rdof[k]); fs.push_back(1); }
  // signal rows: forward entry i uses row (frc_rows == 1 ? 0 : i); adjoint k uses frow_f + k.
  // make_run_buffers maps entry i -> row i unless rows == 1, so give it per-entry rows.
  const int32_t ptot = static_cast<int32_t>(fi.size());
  std::vector<double> Fdual(static_cast<size_t>(rows_dual) * C, 0.0);
  std::vector<double> Fentry(static_cast<size_t>(ptot) * C, 0.0);  // one row per entry
  const int64_t nblk = (n + kRowsSingle - 1) / kRowsSingle;
  RunBuffers R;
  make_run_buffers<T>(R, H, kRowsSingle, nblk, 2, fi.data(), fs.data(), ptot, Fentry.data(),
                      ptot, static_cast<int32_t>(C), nullptr, 0);
  const BlockLists L = R.lists(H);
  T* wI = R.wI.as<T>();
  const IfaceRows<T> IR = iface_rows<T>(H, wI);
  const T* falpha = R.falpha.as<T>();
  DevBuf lam(sizeof(T) * static_cast<size_t>(n) * C);
  DevBuf ubuf(sizeof(T) * static_cast<size_t>(n) * 2 * 3);
  DevBuf zero(sizeof(T) * static_cast<size_t>(n) * 2);
  DevBuf dgK(sizeof(double) * G.n_el), dgM(sizeof(double) * G.n_el);
  BFE_HIP(hipMemsetAsync(ubuf.p, 0, ubuf.bytes, st));
  BFE_HIP(hipMemsetAsync(zero.p, 0, zero.bytes, st));
  BFE_HIP(hipMemsetAsync(dgK.p, 0, dgK.bytes, st));
  BFE_HIP(hipMemsetAsync(dgM.p, 0, dgM.bytes, st));
  T *u = ubuf.as<T>(), *u1 = u + 2 * n, *u2 = u + 4 * n;
  const T dtT = static_cast<T>(H.dt);
  int64_t launches = 0;
  std::vector<T> Fj(static_cast<size_t>(ptot) * C);
  KernelSampler sampler(H, 0, 0);

  BFE_HIP(hipEventRecord(H.ev0, st));
  for (int64_t k = nseg - 1; k >= 0; --k) {
    const int64_t La = sg.len(k);
    const int64_t Lf = k >= 1 ? sg.len(k - 1) : 0;
    const int64_t a_end = sg.start[k + 1];        // adjoint walks t = a_end-1 .. start[k]
    const int64_t f_beg = k >= 1 ? sg.start[k - 1] : 0;
    T* unext = rec.useg[1 - rec.cur].as<T>();
    if (k >= 1) {  // forward column restarts from checkpoint k-1
      const T* ck = rec.ckpt.as<T>() + (k - 1) * 2 * n;
      BFE_HIP(hipMemcpy2DAsync(u1, 2 * sizeof(T), ck, sizeof(T), sizeof(T), n,
                               hipMemcpyDeviceToDevice, st));
      BFE_HIP(hipMemcpy2DAsync(u2, 2 * sizeof(T), ck + n, sizeof(T), sizeof(T), n,
                               hipMemcpyDeviceToDevice, st));
    }
    const int64_t steps = std::max(La, Lf);
    // signals of this sweep, one row per forcing entry, column = sweep step
    for (int64_t j = 0; j < steps; ++j) {
      for (int32_t i = 0; i < pf; ++i)
        Fj[j * ptot + i] = (j < Lf) ? static_cast<T>(rec.F[(f_beg + j) * frow_f + (frow_f == 1 ? 0 : i)])
                                    : T(0);
      for (int32_t r = 0; r < qa; ++r)
        Fj[j * ptot + pf + r] = (j < La) ? static_cast<T>(Fadj[(a_end - 1 - j) * qa + r]) : T(0);
    }
    BFE_HIP(hipMemcpyAsync(R.F.p, Fj.data(), sizeof(T) * static_cast<size_t>(ptot) * steps,
                           hipMemcpyHostToDevice, st));
    for (int64_t j = 0; j < steps; ++j) {
      T* s0 = j < Lf ? unext + j * n : nullptr;
      T* s1 = j < La ? lam.as<T>() + (La - 1 - j) * n : nullptr;
      launch_step<T>(H, 2, nblk, u1, u2, u, L, falpha, R.F.as<T>(), ptot, static_cast<int32_t>(j),
                     0, nullptr, nullptr, 0, dtT, IR, s0, s1);
      ++launches;
      if (H.nI > 0 && H.has_xinv) {
        launch_iface<T>(H, 2, wI, u);
        ++launches;
        // the interface solve rewrote u on interface rows: refresh the stores
        if (s0) BFE_HIP(hipMemcpy2DAsync(s0, sizeof(T), u, 2 * sizeof(T), sizeof(T), n,
                                         hipMemcpyDeviceToDevice, st));
        if (s1) BFE_HIP(hipMemcpy2DAsync(s1, sizeof(T), u + 1, 2 * sizeof(T), sizeof(T), n,
                                         hipMemcpyDeviceToDevice, st));
      }
      T* tmp = u2; u2 = u1; u1 = u; u = tmp;
    }
    // gradient contributions of segment k (states before it: checkpoint k)
    const T* ck = rec.ckpt.as<T>() + k * 2 * n;
    launch_grad<T>(H, G, rec.useg[rec.cur].as<T>(), lam.as<T>(), static_cast<int32_t>(La), ck,
                   ck + n, dgK.as<double>(), dgM.as<double>());
    ++launches;
    rec.cur = 1 - rec.cur;
  }
  finish_timing(H, sampler, n_t - rec.t0, launches);
  BFE_HIP(hipMemcpy(g_stiff, dgK.p, dgK.bytes, hipMemcpyDeviceToHost));
  BFE_HIP(hipMemcpy(g_dens, dgM.p, dgM.bytes, hipMemcpyDeviceToHost));
}

template <typename T, bool D, int R, int G>
void launch_multi_r(const bfe_plan_s& H, int64_t nblk, int32_t rpb, int64_t n, int32_t S,
                    const T* u1, const T* u2, T* u, const BlockLists& L, const T* falpha,
                    const T* F, int32_t F_rows, int32_t t, int32_t q, T* hp,
                    const IfaceRows<T>& IR) {
  const GroupOp& g = H.grp[G == 1 ? 0 : G == 2 ? 1 : G == 4 ? 2 : 3];
  const DevBuf &o = g.off, &c = g.col, &v = g.val;
  hipLaunchKernelGGL((k_step_multi<T, D, R, G>), dim3(nblk), dim3(64 * rpb), 0, H.stream, n, S,
                     nblk, H.xcd_remap, o.as<int64_t>(), c.as<int32_t>(), v.as<T>(),
                     D ? H.beta.as<T>() : nullptr, u1, u2, u, L, falpha, F, F_rows, t, q, hp, IR);
}

// Rows per wavefront R and rows per gather group G (G <= R, both powers of 2).
template <typename T, bool D>
void launch_multi_d(const bfe_plan_s& H, int64_t nblk, int32_t rpb, int32_t rpw, int32_t G,
                    int64_t n, int32_t S, const T* u1, const T* u2, T* u, const BlockLists& L,
                    const T* falpha, const T* F, int32_t F_rows, int32_t t, int32_t q, T* hp,
                    const IfaceRows<T>& IR) {
#define BFE_MULTI(R_, G_)                                                                   \
  case R_ * 16 + G_:                                                                        \
    launch_multi_r<T, D, R_, G_>(H, nblk, rpb, n, S, u1, u2, u, L, falpha, F, F_rows, t, q, \
                                 hp, IR);                                                   \
    break;
  switch (rpw * 16 + G) {
    BFE_MULTI(1, 1)
    BFE_MULTI(2, 1) BFE_MULTI(2, 2)
    BFE_MULTI(4, 1) BFE_MULTI(4, 2) BFE_MULTI(4, 4)
    BFE_MULTI(8, 1) BFE_MULTI(8, 2) BFE_MULTI(8, 4) BFE_MULTI(8, 8)
    BFE_MULTI(16, 1) BFE_MULTI(16, 2) BFE_MULTI(16, 4) BFE_MULTI(16, 8)
    default: throw std::invalid_argument("rows per wavefront must be 1, 2, 4, 8 or 16");
  }
#undef BFE_MULTI
}

template <typename T>
void launch_multi(const bfe_plan_s& H, int64_t nblk, int32_t rpb, int32_t rpw, int32_t G,
                  int64_t n, int32_t S, const T* u1, const T* u2, T* u, const BlockLists& L,
                  const T* falpha, const T* F, int32_t F_rows, int32_t t, int32_t q, T* hp,
                  const IfaceRows<T>& IR) {
  if (H.damped)
    launch_multi_d<T, true>(H, nblk, rpb, rpw, G, n, S, u1, u2, u, L, falpha, F, F_rows, t, q, hp, IR);
  else
    launch_multi_d<T, false>(H, nblk, rpb, rpw, G, n, S, u1, u2, u, L, falpha, F, F_rows, t, q, hp, IR);
}

// Per-group lists of one batched run on the element operator (group-local
// rows, grouped per workgroup of k_step_ebe_multi: 4 wavefronts of npw nodes).
struct EbeMultiLists {
  DevBuf fptr, frow, fsrc, ffrow, falpha, mptr, mrow, mk, iptr, irow, ik;
  BlockLists lists() const {
    BlockLists L;
    L.fptr = fptr.as<int32_t>(); L.f_row = frow.as<int32_t>(); L.f_src = fsrc.as<int32_t>();
    L.f_frow = ffrow.as<int32_t>();
    L.mptr = mptr.as<int32_t>(); L.m_row = mrow.as<int32_t>(); L.m_k = mk.as<int32_t>();
    L.iptr = iptr.as<int32_t>(); L.i_row = irow.as<int32_t>(); L.i_k = ik.as<int32_t>();
    return L;
  }
};

constexpr int kMultiWaves = kBlock / 64;  // wavefronts per workgroup of k_step_ebe_multi

inline int64_t ebe_multi_blocks(const EbeGroup& Gr, int32_t npw) {
  const int64_t per = static_cast<int64_t>(kMultiWaves) * npw;
  return (Gr.n_nodes + per - 1) / per;
}

// Whether the batched march can run on the element operator: every group
// has a pattern kernel.
bool ebe_multi_ok(const bfe_plan_s& H) {
  if (!H.ebe || !H.ebe_kernel) return false;
  for (const auto& g : H.ebe->groups)
    if (g.pattern == 0 || !has_pattern_kernel(g.pattern, g.dpn)) return false;
  return true;
}

// Forcing (for one (dof, source) the last entry wins, as make_run_buffers),
// monitor and interface lists of every group for one batched run.
template <typename T>
std::vector<EbeMultiLists> ebe_multi_lists(const bfe_plan_s& H, const EbeOp& E, int32_t npw,
                                           const int32_t* frc_idx, const int32_t* frc_src,
                                           int32_t p, int32_t frc_rows, const int32_t* hist_idx,
                                           int32_t q) {
  std::vector<std::pair<std::pair<int32_t, int32_t>, int32_t>> key;
  for (int32_t i = 0; i < p; ++i) key.push_back({{frc_idx[i], frc_src[i]}, i});
  std::stable_sort(key.begin(), key.end(), [](auto& a, auto& b) { return a.first < b.first; });
  const size_t ng = E.groups.size();
  std::vector<std::vector<int32_t>> frows(ng), fsrc(ng), ffrow(ng), fidx(ng), mrows(ng), mks(ng),
      irows(ng), iks(ng);
  std::vector<std::vector<double>> fal(ng);
  for (size_t i = 0; i < key.size(); ++i)
    if (i + 1 == key.size() || key[i + 1].first != key[i].first) {
      const int32_t r = key[i].first.first;
      const int g = E.row_group[r];
      if (g < 0) throw std::invalid_argument("forcing on a DOF that no element group owns");
      frows[g].push_back(E.row_local[r]);
      fsrc[g].push_back(key[i].first.second);
      ffrow[g].push_back(frc_rows == 1 ? 0 : key[i].second);
      fal[g].push_back(H.alpha[r]);
    }
  for (int32_t k = 0; k < q; ++k) {
    const int32_t r = hist_idx[k];
    const int g = E.row_group[r];
    if (g < 0) throw std::invalid_argument("monitor on a DOF that no element group owns");
    mrows[g].push_back(E.row_local[r]);
    mks[g].push_back(k);
  }
  for (size_t k = 0; k < H.iface_host.size(); ++k) {
    const int32_t r = H.iface_host[k];
    const int g = E.row_group[r];
    if (g < 0) throw std::invalid_argument("interface row that no element group owns");
    irows[g].push_back(E.row_local[r]);
    iks[g].push_back(static_cast<int32_t>(k));
  }
  std::vector<EbeMultiLists> out(ng);
  for (size_t g = 0; g < ng; ++g) {
    const EbeGroup& Gr = E.groups[g];
    const int64_t nb = ebe_multi_blocks(Gr, npw);
    const int64_t rpb = static_cast<int64_t>(kMultiWaves) * npw * Gr.dpn;
    std::vector<int32_t> idx(frows[g].size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = static_cast<int32_t>(i);
    Grouped gf = group_by_block(nb, rpb, frows[g], &fsrc[g], &ffrow[g]);
    Grouped gfi = group_by_block(nb, rpb, frows[g], &idx, nullptr);
    std::vector<double> fa(gfi.a.size());
    for (size_t i = 0; i < fa.size(); ++i) fa[i] = fal[g][gfi.a[i]];
    Grouped gm = group_by_block(nb, rpb, mrows[g], &mks[g], nullptr);
    Grouped gi = group_by_block(nb, rpb, irows[g], &iks[g], nullptr);
    EbeMultiLists& R = out[g];
    R.fptr = upload(gf.ptr); R.frow = upload(gf.row); R.fsrc = upload(gf.a);
    R.ffrow = upload(gf.b); R.falpha = upload(cast_vec<T>(fa));
    R.mptr = upload(gm.ptr); R.mrow = upload(gm.row); R.mk = upload(gm.a);
    R.iptr = upload(gi.ptr); R.irow = upload(gi.row); R.ik = upload(gi.a);
  }
  return out;
}

template <typename T, bool D, typename PT, int DPN>
void launch_ebe_multi_pat(const bfe_plan_s& H, const EbeOp& E, const EbeGroup& Gr,
                          const EbeMultiLists& ML, int32_t npw, int32_t S, const T* u1,
                          const T* u2, T* u, const T* F, int32_t F_rows, int32_t t, int32_t q,
                          T* hp, const IfaceRows<T>& IR, bool node_blocks) {
  ClassOffsets<PT::NC> co;
  for (int j = 0; j < PT::NC; ++j) co.off[j] = static_cast<int32_t>(Gr.doff[j]);
  AnchorSlots<PT::NS> an{};
  const bool anchored = PT::NS <= kAnchorMaxSlots && Gr.anchor.bytes > 0;
  if (anchored) {
    for (int k = 0; k < PT::NS; ++k)
      an.base[k] = static_cast<int32_t>(static_cast<int64_t>(Gr.anc_row[k]) * Gr.n_nodes +
                                        Gr.anc_noff[k]);
    an.len = static_cast<int32_t>(Gr.anchor.bytes / sizeof(T));
  }
  const int32_t slide = (H.multi_pipeline && offsets_follow_next<PT>(Gr.doff, DPN)) ? 1 : 0;
  const int64_t nblk = ebe_multi_blocks(Gr, npw);
  const BlockLists L = ML.lists();
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(kBlock), 0, H.stream, Gr.n_nodes, nblk, npw, S,
                       slide, H.n, Gr.uniform ? nullptr : Gr.dof0.as<int32_t>(), co, an,
                       Gr.pmask.as<uint32_t>(), Gr.smask.as<uint32_t>(), Gr.nbr.as<int32_t>(),
                       Gr.adj.as<int32_t>(), Gr.scale.as<T>(),
                       anchored ? Gr.anchor.as<T>() : static_cast<const T*>(nullptr),
                       Gr.srows_c.as<T>(), Gr.stencil.as<T>(), E.alpha_of<T>(Gr),
                       D ? E.beta_of<T>(Gr) : nullptr, u1, u2, u, L, ML.falpha.as<T>(), F,
                       F_rows, t, q, hp, IR, Gr.coef_node, Gr.nblk.as<T>());
  };
  if constexpr (PT::NPE == 3) {
    if (node_blocks) {
      go(k_step_ebe_multi<T, D, PT, DPN, false, false, true>);
      return;
    }
  }
  if constexpr (PT::NS <= kAnchorMaxSlots) {
    if (anchored) {
      go(k_step_ebe_multi<T, D, PT, DPN, true>);
      return;
    }
  }
  if (std::is_same<PT, PatKuhn24>::value && Gr.sparse)
    go(k_step_ebe_multi<T, D, PT, DPN, false, true>);
  else
    go(k_step_ebe_multi<T, D, PT, DPN, false, false>);
}

// One batched step of every group of the element operator.
template <typename T>
void launch_step_ebe_multi(const bfe_plan_s& H, const EbeOp& E,
                           const std::vector<EbeMultiLists>& ML, int32_t npw, int32_t S,
                           const T* u1, const T* u2, T* u, const T* F, int32_t F_rows, int32_t t,
                           int32_t q, T* hp, const IfaceRows<T>& IR, bool node_blocks) {
  for (size_t g = 0; g < E.groups.size(); ++g) {
    const EbeGroup& Gr = E.groups[g];
#define BFE_EM(D_, PT_, DPN_)                                                                  \
  launch_ebe_multi_pat<T, D_, PT_, DPN_>(H, E, Gr, ML[g], npw, S, u1, u2, u, F, F_rows, t, q, hp, \
                                         IR, node_blocks)
#define BFE_EM_D(PT_, DPN_) \
  if (H.damped) BFE_EM(true, PT_, DPN_); else BFE_EM(false, PT_, DPN_);
    switch (Gr.pattern * 10 + Gr.dpn) {
      case PatKuhn24::ID * 10 + 3: BFE_EM_D(PatKuhn24, 3) break;
      case PatTriRect::ID * 10 + 1: BFE_EM_D(PatTriRect, 1) break;
      case PatTriRect::ID * 10 + 2: BFE_EM_D(PatTriRect, 2) break;
      case PatTriRect::ID * 10 + 3: BFE_EM_D(PatTriRect, 3) break;
      case PatTriIso::ID * 10 + 1: BFE_EM_D(PatTriIso, 1) break;
      case PatTriIso::ID * 10 + 2: BFE_EM_D(PatTriIso, 2) break;
      case PatTriIso::ID * 10 + 3: BFE_EM_D(PatTriIso, 3) break;
      default: throw std::invalid_argument("no multi-source kernel for this element group");
    }
#undef BFE_EM_D
#undef BFE_EM
  }
}

// Whether every group of the element operator can take the node-block form
// (triangle patterns).
bool node_blocks_ok(const bfe_plan_s& H) {
  if (!ebe_multi_ok(H)) return false;
  for (const auto& g : H.ebe->groups)
    if (g.pattern != PatTriRect::ID && g.pattern != PatTriIso::ID) return false;
  return true;
}

// Builds the node-block form of every group that does not hold it yet.
template <typename T, typename PT, int DPN>
void build_node_blocks_pat(const bfe_plan_s& H, EbeGroup& Gr) {
  AnchorSlots<PT::NS> an{};
  const bool anchored = PT::NS <= kAnchorMaxSlots && Gr.anchor.bytes > 0;
  if (anchored) {
    for (int k = 0; k < PT::NS; ++k)
      an.base[k] = static_cast<int32_t>(static_cast<int64_t>(Gr.anc_row[k]) * Gr.n_nodes +
                                        Gr.anc_noff[k]);
    an.len = static_cast<int32_t>(Gr.anchor.bytes / sizeof(T));
  }
  Gr.nblk = DevBuf(sizeof(T) * static_cast<size_t>(Gr.n_nodes) * PT::NC * DPN * DPN);
  const dim3 grid((Gr.n_nodes + kBlock - 1) / kBlock), blk(kBlock);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, blk, 0, H.stream, Gr.n_nodes, an, Gr.pmask.as<uint32_t>(),
                       Gr.smask.as<uint32_t>(), Gr.adj.as<int32_t>(), Gr.scale.as<T>(),
                       anchored ? Gr.anchor.as<T>() : static_cast<const T*>(nullptr),
                       Gr.srows_c.as<T>(), Gr.nblk.as<T>());
  };
  if (anchored) go(k_node_blocks<T, PT, DPN, true>);
  else go(k_node_blocks<T, PT, DPN, false>);
  BFE_HIP(hipGetLastError());
}

template <typename T>
void ensure_node_blocks(const bfe_plan_s& H, EbeOp& E) {
  for (auto& Gr : E.groups) {
    if (Gr.nblk.bytes > 0) continue;
    switch (Gr.pattern * 10 + Gr.dpn) {
      case PatTriRect::ID * 10 + 1: build_node_blocks_pat<T, PatTriRect, 1>(H, Gr); break;
      case PatTriRect::ID * 10 + 2: build_node_blocks_pat<T, PatTriRect, 2>(H, Gr); break;
      case PatTriRect::ID * 10 + 3: build_node_blocks_pat<T, PatTriRect, 3>(H, Gr); break;
      case PatTriIso::ID * 10 + 1: build_node_blocks_pat<T, PatTriIso, 1>(H, Gr); break;
      case PatTriIso::ID * 10 + 2: build_node_blocks_pat<T, PatTriIso, 2>(H, Gr); break;
      case PatTriIso::ID * 10 + 3: build_node_blocks_pat<T, PatTriIso, 3>(H, Gr); break;
      default: throw std::invalid_argument("no node-block form for this element group");
    }
  }
}

template <typename T>
void run_batched(bfe_plan_s& H, int32_t n_t, int32_t n_src, const int32_t* frc_idx,
                 const int32_t* frc_src, int32_t p, const double* frcs, int32_t frc_rows,
                 const int32_t* hist_idx, int32_t q, double* hist_out,
                 int32_t* ti_start_out, AdjointRecord* rec = nullptr, int32_t ckpt_every = 0) {
  const int64_t n = H.n;
  validate_run(n, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out);
  if (n_src < 1 || n_src > 64) throw std::invalid_argument("n_src must be in [1, 64]");
  check_indices(frc_src, p, n_src, "forcing source");
  const int32_t S = n_src;
  H.stats = bfe_run_stats{};
  H.stats.columns = S;
  const int64_t t0 = p > 0 ? first_forced_step(frcs, frc_rows, n_t) : -1;
  if (ti_start_out) *ti_start_out = t0 < 0 ? n_t + 1 : static_cast<int32_t>(t0 + 1);
  if (rec) {
    *rec = AdjointRecord{};
    rec->n_t = n_t;
    rec->p = p;
    rec->frc_rows = frc_rows;
    rec->q = q;
    rec->t0 = t0;
    rec->batched = true;
    rec->S = S;
    rec->frc_idx.assign(frc_idx, frc_idx + p);
    rec->frc_src.assign(frc_src, frc_src + p);
    rec->hist_idx.assign(hist_idx, hist_idx + q);
    if (p > 0) rec->F.assign(frcs, frcs + static_cast<size_t>(frc_rows) * n_t);
    rec->valid = true;
  }
  if (t0 < 0) {
    if (q > 0) std::memset(hist_out, 0, sizeof(double) * static_cast<size_t>(q) * S * n_t);
    return;
  }

  BFE_HIP(hipSetDevice(H.device));
  hipStream_t st = H.stream;
  const int32_t rpb = H.multi_rows;
  const int32_t rpw = H.multi_rpw;
  // the element operator's multi-source product when every group has a
  // pattern kernel, else the row-grouped CSR operator (a checkpointed run
  // stays on the CSR operator that the adjoint's recomputation uses)
  const bool nbf = H.multi_ebe == BFE_OPERATOR_NODE && rec == nullptr && node_blocks_ok(H);
  const EbeOp* E = ((H.multi_ebe == BFE_OPERATOR_EBE && rec == nullptr && ebe_multi_ok(H)) || nbf)
                       ? H.ebe.get()
                       : nullptr;
  H.last_multi_op = nbf ? BFE_OPERATOR_NODE : E ? BFE_OPERATOR_EBE : BFE_OPERATOR_CSR;
  if (nbf) ensure_node_blocks<T>(H, *H.ebe);
  const int32_t G = E ? 0 : resolve_group<T>(H, rpw);
  H.last_group = G;
  const int64_t rows_per_wg = static_cast<int64_t>(rpb) * rpw;
  const int64_t nblk = (n + rows_per_wg - 1) / rows_per_wg;
  RunBuffers R;
  make_run_buffers<T>(R, H, rows_per_wg, nblk, S, frc_idx, frc_src, p, frcs, frc_rows, n_t,
                      hist_idx, q);
  BlockLists L = R.lists(H);
  // interface rows grouped per workgroup of this launch shape
  std::vector<int32_t> ks(H.iface_host.size());
  for (size_t k = 0; k < ks.size(); ++k) ks[k] = static_cast<int32_t>(k);
  Grouped gi = group_by_block(nblk, rows_per_wg, H.iface_host, &ks, nullptr);
  DevBuf gi_ptr = upload(gi.ptr), gi_row = upload(gi.row), gi_k = upload(gi.a);
  L.iptr = gi_ptr.as<int32_t>();
  L.i_row = gi_row.as<int32_t>();
  L.i_k = gi_k.as<int32_t>();
  std::vector<EbeMultiLists> ML;
  if (E)
    ML = ebe_multi_lists<T>(H, *E, H.multi_npw, frc_idx, frc_src, p, frc_rows, hist_idx, q);
  DevBuf dhist(sizeof(T) * static_cast<size_t>(q) * S * n_t);
  DevBuf ubuf(sizeof(T) * static_cast<size_t>(n) * S * 3);
  BFE_HIP(hipMemsetAsync(ubuf.p, 0, ubuf.bytes, st));
  if (dhist.bytes) BFE_HIP(hipMemsetAsync(dhist.p, 0, dhist.bytes, st));
  const int64_t ns = n * S;
  T *u = ubuf.as<T>(), *u1 = u + ns, *u2 = u + 2 * ns;
  T* hist = dhist.as<T>();
  T* wI = R.wI.as<T>();
  const IfaceRows<T> IR = iface_rows<T>(H, wI);
  const T* falpha = R.falpha.as<T>();
  const T* F = R.F.as<T>();
  const int64_t per_t = static_cast<int64_t>(q) * S;
  int64_t launches = 0;
  int64_t next_seg = 0;
  if (rec) {
    rec->seg = make_segments(t0, n_t, ckpt_every);
    rec->ckpt = DevBuf(sizeof(T) * static_cast<size_t>(ns) * 2 * rec->seg.count());
    rec->useg[0] = DevBuf(sizeof(T) * static_cast<size_t>(ns) * rec->seg.C);
  }
  KernelSampler sampler(H, t0, n_t - t0);

  BFE_HIP(hipEventRecord(H.ev0, st));
  for (int64_t t = t0; t < n_t; ++t) {
    const int64_t prev = t - 1;
    T* hp = (q > 0 && prev >= t0) ? hist + prev * per_t : nullptr;
    const int32_t ti = static_cast<int32_t>(t);
    if (rec && next_seg < rec->seg.count() && t == rec->seg.start[next_seg]) {
      // checkpoint: the two states the step at t starts from
      T* ck = rec->ckpt.as<T>() + next_seg * 2 * ns;
      BFE_HIP(hipMemcpyAsync(ck, u1, sizeof(T) * ns, hipMemcpyDeviceToDevice, st));
      BFE_HIP(hipMemcpyAsync(ck + ns, u2, sizeof(T) * ns, hipMemcpyDeviceToDevice, st));
      ++next_seg;
    }
    sampler.before(t);
    if (E) {
      launch_step_ebe_multi<T>(H, *E, ML, H.multi_npw, S, u1, u2, u, F, frc_rows, ti, q, hp, IR,
                               nbf);
      launches += static_cast<int64_t>(E->groups.size());
    } else {
      launch_multi<T>(H, nblk, rpb, rpw, G, n, S, u1, u2, u, L, falpha, F, frc_rows, ti, q, hp,
                      IR);
      ++launches;
    }
    sampler.after(t);
    if (H.nI > 0 && H.has_xinv) {
      launch_iface<T>(H, S, wI, u);
      ++launches;
    }
    if (rec) {  // every state of the last segment
      const int64_t last = rec->seg.start[rec->seg.count() - 1];
      if (t >= last)
        BFE_HIP(hipMemcpyAsync(rec->useg[0].as<T>() + (t - last) * ns, u, sizeof(T) * ns,
                               hipMemcpyDeviceToDevice, st));
    }
    T* tmp = u2; u2 = u1; u1 = u; u = tmp;
  }
  if (q > 0) {
    hipLaunchKernelGGL((k_snap<T>), dim3((per_t + 255) / 256), dim3(256), 0, st, n, S, u1, u2, q,
                       R.hidx.as<int32_t>(), hist + (n_t - 1) * per_t, static_cast<T*>(nullptr),
                       0, static_cast<T>(H.dt));
    ++launches;
  }
  finish_timing(H, sampler, n_t - t0, launches);
  if (q > 0) H.stats.nonfinite = copy_out<T>(dhist, hist_out, static_cast<size_t>(per_t) * n_t);
}

// The reverse sweep of a batched checkpointed run (S sources): one
// multi-source march of 2S columns per step, columns 0..S-1 recomputing
// segment k-1 from its checkpoint and columns S..2S-1 marching the adjoint
// over segment k (the sources' receiver residuals as forcing), so the
// operator is read once for both. Each sweep writes its states straight
// into a C x n x 2S buffer (no copies); two buffers alternate, because the
// gradient of segment k pairs the forward states recomputed in sweep k + 1
// with the adjoint states of sweep k. The gradients are summed over the
// sources by k_grad_multi.
template <typename T>
void run_adjoint_batched(bfe_plan_s& H, AdjointRecord& rec, const GradElements& G,
                         const double* hist_grad, double* g_stiff, double* g_dens) {
  const int64_t n = H.n;
  const int32_t S = rec.S, S2 = 2 * rec.S;
  if (S2 > 64) throw std::invalid_argument("a batched adjoint takes at most 32 sources");
  std::memset(g_stiff, 0, sizeof(double) * G.n_el);
  std::memset(g_dens, 0, sizeof(double) * G.n_el);
  H.stats = bfe_run_stats{};
  H.stats.columns = S2;
  if (rec.t0 < 0) return;
  BFE_HIP(hipSetDevice(H.device));
  hipStream_t st = H.stream;
  const int32_t n_t = rec.n_t;
  const Segments& sg = rec.seg;
  const int64_t nseg = sg.count(), C = sg.C;
  const double dt2 = H.dt * H.dt;
  const int32_t q = rec.q;

  // adjoint forcing: receivers with duplicate DOFs summed, per source, / dt^2
  std::vector<int32_t> rdof;
  std::vector<int32_t> rrow(q);
  {
    std::vector<std::pair<int32_t, int32_t>> kk;
    for (int32_t k = 0; k < q; ++k) kk.push_back({rec.hist_idx[k], k});
    std::sort(kk.begin(), kk.end());
    for (auto& e : kk) {
      if (rdof.empty() || rdof.back() != e.first) rdof.push_back(e.first);
      rrow[e.second] = static_cast<int32_t>(rdof.size()) - 1;
    }
  }
  const int32_t qa = static_cast<int32_t>(rdof.size());
  // Fadj (t, receiver, source); hist_grad element (k, s, t) at k + q (s + S t)
  std::vector<double> Fadj(static_cast<size_t>(qa) * S * n_t, 0.0);
  for (int64_t t = 0; t < n_t; ++t)
    for (int32_t s = 0; s < S; ++s)
      for (int32_t k = 0; k < q; ++k)
        Fadj[(t * S + s) * qa + rrow[k]] += hist_grad[k + static_cast<int64_t>(q) * (s + S * t)] / dt2;

  // forcing entries of the 2S-column march: the forward entries in their
  // source's column, then one adjoint entry per (receiver, source)
  const int32_t pf = rec.p, frow_f = rec.frc_rows;
  std::vector<int32_t> fi, fs;
  for (int32_t i = 0; i < pf; ++i) { fi.push_back(rec.frc_idx[i]); fs.push_back(rec.frc_src[i]); }
  for (int32_t s = 0; s < S; ++s)
    for (int32_t k = 0; k < qa; ++k) { fi.push_back(rdof[k]); fs.push_back(S + s); }
  const int32_t ptot = static_cast<int32_t>(fi.size());
  const int32_t rpb = H.multi_rows, rpw = H.multi_rpw;
  const int32_t Gg = resolve_group<T>(H, rpw);
  const int64_t rows_per_wg = static_cast<int64_t>(rpb) * rpw;
  const int64_t nblk = (n + rows_per_wg - 1) / rows_per_wg;
  std::vector<double> Fentry(static_cast<size_t>(ptot) * C, 0.0);  // one row per entry
  RunBuffers R;
  make_run_buffers<T>(R, H, rows_per_wg, nblk, S2, fi.data(), fs.data(), ptot, Fentry.data(),
                      ptot, static_cast<int32_t>(C), nullptr, 0);
  BlockLists L = R.lists(H);
  std::vector<int32_t> ks(H.iface_host.size());
  for (size_t k = 0; k < ks.size(); ++k) ks[k] = static_cast<int32_t>(k);
  Grouped gi = group_by_block(nblk, rows_per_wg, H.iface_host, &ks, nullptr);
  DevBuf gi_ptr = upload(gi.ptr), gi_row = upload(gi.row), gi_k = upload(gi.a);
  L.iptr = gi_ptr.as<int32_t>();
  L.i_row = gi_row.as<int32_t>();
  L.i_k = gi_k.as<int32_t>();
  T* wI = R.wI.as<T>();
  const IfaceRows<T> IR = iface_rows<T>(H, wI);
  const T* falpha = R.falpha.as<T>();
  const int64_t nS = n * S, nS2 = n * S2;
  DevBuf buf[2], init[2];
  for (int b = 0; b < 2; ++b) {
    buf[b] = DevBuf(sizeof(T) * static_cast<size_t>(nS2) * C);
    init[b] = DevBuf(sizeof(T) * static_cast<size_t>(nS2) * 2);  // the two states before a sweep
  }
  DevBuf dgK(sizeof(double) * G.n_el), dgM(sizeof(double) * G.n_el);
  BFE_HIP(hipMemsetAsync(dgK.p, 0, dgK.bytes, st));
  BFE_HIP(hipMemsetAsync(dgM.p, 0, dgM.bytes, st));
  const bool pairs = H.grad_pairs && G.pair_gd > 0;
  DevBuf pacc(pairs ? sizeof(double) * static_cast<size_t>(G.n_groups) * G.pair_stride : 0);
  if (pairs) BFE_HIP(hipMemsetAsync(pacc.p, 0, pacc.bytes, st));
  int64_t launches = 0;
  std::vector<T> Fj(static_cast<size_t>(ptot) * C);
  KernelSampler sampler(H, 0, 0);
  const size_t col_bytes = sizeof(T) * S;  // one row's S columns

  BFE_HIP(hipEventRecord(H.ev0, st));
  const T* last1 = nullptr;  // the adjoint states the previous sweep ended with
  const T* last2 = nullptr;
  for (int64_t k = nseg - 1; k >= 0; --k) {
    const int64_t La = sg.len(k);
    const int64_t Lf = k >= 1 ? sg.len(k - 1) : 0;
    const int64_t a_end = sg.start[k + 1];        // adjoint walks t = a_end-1 .. start[k]
    const int64_t f_beg = k >= 1 ? sg.start[k - 1] : 0;
    T* bk = buf[k & 1].as<T>();
    T* i1 = init[k & 1].as<T>();
    T* i2 = i1 + nS2;
    // the forward columns restart from checkpoint k-1 (zero before the first
    // segment); the adjoint columns continue from the previous sweep
    if (k >= 1) {
      const T* ck = rec.ckpt.as<T>() + (k - 1) * 2 * nS;
      BFE_HIP(hipMemcpy2DAsync(i1, sizeof(T) * S2, ck, col_bytes, col_bytes, n,
                               hipMemcpyDeviceToDevice, st));
      BFE_HIP(hipMemcpy2DAsync(i2, sizeof(T) * S2, ck + nS, col_bytes, col_bytes, n,
                               hipMemcpyDeviceToDevice, st));
    } else {
      BFE_HIP(hipMemset2DAsync(i1, sizeof(T) * S2, 0, col_bytes, n, st));
      BFE_HIP(hipMemset2DAsync(i2, sizeof(T) * S2, 0, col_bytes, n, st));
    }
    if (last1 != nullptr) {
      BFE_HIP(hipMemcpy2DAsync(i1 + S, sizeof(T) * S2, last1 + S, sizeof(T) * S2, col_bytes, n,
                               hipMemcpyDeviceToDevice, st));
      BFE_HIP(hipMemcpy2DAsync(i2 + S, sizeof(T) * S2, last2 + S, sizeof(T) * S2, col_bytes, n,
                               hipMemcpyDeviceToDevice, st));
    } else {
      BFE_HIP(hipMemset2DAsync(i1 + S, sizeof(T) * S2, 0, col_bytes, n, st));
      BFE_HIP(hipMemset2DAsync(i2 + S, sizeof(T) * S2, 0, col_bytes, n, st));
    }
    const int64_t steps = std::max(La, Lf);
    // signals of this sweep, one row per forcing entry, column = sweep step
    for (int64_t j = 0; j < steps; ++j) {
      for (int32_t i = 0; i < pf; ++i)
        Fj[j * ptot + i] = (j < Lf) ? static_cast<T>(rec.F[(f_beg + j) * frow_f + (frow_f == 1 ? 0 : i)])
                                    : T(0);
      for (int32_t s = 0; s < S; ++s)
        for (int32_t r = 0; r < qa; ++r)
          Fj[j * ptot + pf + s * qa + r] =
              (j < La) ? static_cast<T>(Fadj[((a_end - 1 - j) * S + s) * qa + r]) : T(0);
    }
    BFE_HIP(hipMemcpyAsync(R.F.p, Fj.data(), sizeof(T) * static_cast<size_t>(ptot) * steps,
                           hipMemcpyHostToDevice, st));
    for (int64_t j = 0; j < steps; ++j) {
      T* u = bk + j * nS2;
      const T* u1 = j >= 1 ? bk + (j - 1) * nS2 : i1;
      const T* u2 = j >= 2 ? bk + (j - 2) * nS2 : (j == 1 ? i1 : i2);
      launch_multi<T>(H, nblk, rpb, rpw, Gg, n, S2, u1, u2, u, L, falpha, R.F.as<T>(), ptot,
                      static_cast<int32_t>(j), 0, nullptr, IR);
      ++launches;
      if (H.nI > 0 && H.has_xinv) {
        launch_iface<T>(H, S2, wI, u);
        ++launches;
      }
    }
    last1 = steps >= 1 ? bk + (steps - 1) * nS2 : i1;
    last2 = steps >= 2 ? bk + (steps - 2) * nS2 : (steps == 1 ? i1 : i2);
    // gradient of segment k: its forward states (the last segment's from the
    // forward run, the others recomputed in sweep k + 1) with the adjoint
    // states of this sweep, stored in reverse time
    const T* ck = rec.ckpt.as<T>() + k * 2 * nS;
    const T* lam = bk + (La - 1) * nS2 + S;
    const T* Useg = k == nseg - 1 ? rec.useg[0].as<T>() : buf[(k + 1) & 1].as<T>();
    const int64_t us = k == nseg - 1 ? nS : nS2, ud = k == nseg - 1 ? S : S2;
    if (pairs)
      launch_pair_acc<T>(H, G, S, Useg, us, ud, lam, -nS2, S2, static_cast<int32_t>(La), ck,
                         ck + nS, pacc.as<double>());
    else
      launch_grad_multi<T>(H, G, S, Useg, us, ud, lam, -nS2, S2, static_cast<int32_t>(La), ck,
                           ck + nS, dgK.as<double>(), dgM.as<double>());
    ++launches;
  }
  if (pairs) {
    launch_pair_contract<T>(H, G, pacc.as<double>(), dgK.as<double>(), dgM.as<double>());
    ++launches;
  }
  finish_timing(H, sampler, n_t - rec.t0, launches);
  BFE_HIP(hipMemcpy(g_stiff, dgK.p, dgK.bytes, hipMemcpyDeviceToHost));
  BFE_HIP(hipMemcpy(g_dens, dgM.p, dgM.bytes, hipMemcpyDeviceToHost));
}

}  // namespace

namespace {

// Builds, checks and attaches the element-by-element operator (see
// bfe_set_element_groups). Throws on malformed input or when the operator
// does not reproduce the CSR step.
void set_element_groups(bfe_plan_s* h, int32_t n_groups, const bfe_element_group* groups,
                        double check_tol) {
  if (h == nullptr) throw std::invalid_argument("NULL handle");
  if (n_groups < 1 || n_groups > 8) throw std::invalid_argument("1 to 8 element groups");
  if (groups == nullptr) throw std::invalid_argument("NULL element groups");
  const int64_t n = h->n;
  auto E = std::make_unique<EbeOp>();
  E->row_group.assign(n, -1);
  E->row_local.assign(n, 0);
  BFE_HIP(hipSetDevice(h->device));
  for (int32_t gi = 0; gi < n_groups; ++gi) {
    const bfe_element_group& in = groups[gi];
    const int64_t nn = in.n_nodes, ne = in.n_el;
    const int32_t dpn = in.dpn, npe = in.npe;
    if (!has_generic_kernel(npe, dpn))
      throw std::invalid_argument("element group: supported (nodes, DOF per node) are "
                                  "(3, 1), (3, 2), (3, 3) and (4, 3)");
    if (nn < 1 || nn > INT32_MAX || ne < 1 || ne > INT32_MAX)
      throw std::invalid_argument("element group: node or element count out of range");
    if (in.n_tmpl < 1) throw std::invalid_argument("element group: no templates");
    if (!in.el_nodes || !in.el_tmpl || !in.el_scale || !in.tmpl)
      throw std::invalid_argument("element group: NULL array");
    if (in.node_dof == nullptr && (n_groups != 1 || nn * dpn != n))
      throw std::invalid_argument("element group without a DOF table: n must equal "
                                  "n_nodes * dof_per_node (node-major numbering)");
    const int KE = npe * dpn;
    for (int64_t e = 0; e < ne; ++e) {
      if (in.el_tmpl[e] < 0 || in.el_tmpl[e] >= in.n_tmpl)
        throw std::invalid_argument("element template index out of range");
      if (!std::isfinite(in.el_scale[e])) throw std::invalid_argument("non-finite element scale");
      for (int l = 0; l < npe; ++l) {
        const int32_t v = in.el_nodes[e * npe + l];
        if (v < 0 || v >= nn) throw std::invalid_argument("element node index out of range");
      }
    }
    for (int64_t i = 0; i < static_cast<int64_t>(in.n_tmpl) * KE * KE; ++i)
      if (!std::isfinite(in.tmpl[i])) throw std::invalid_argument("non-finite template entry");
    // first DOF of every node; every (node, component) owned by one group
    std::vector<int32_t> dof0(nn);
    for (int64_t j = 0; j < nn; ++j) {
      const int64_t d = in.node_dof ? in.node_dof[j] : j * dpn;
      if (d < 0 || d + dpn > n) throw std::invalid_argument("element group DOF out of range");
      dof0[j] = static_cast<int32_t>(d);
      for (int c = 0; c < dpn; ++c) {
        if (E->row_group[d + c] != -1)
          throw std::invalid_argument("a DOF belongs to two element groups");
        E->row_group[d + c] = static_cast<int8_t>(gi);
        E->row_local[d + c] = static_cast<int32_t>(j * dpn + c);
      }
    }
    bfe::EbeLayout Lay = bfe::build_ebe_layout(nn, dpn, ne, npe, in.el_nodes, in.el_tmpl,
                                                in.n_tmpl, in.tmpl, kEbeSlotPad, 64);
    EbeGroup Gr;
    Gr.n_nodes = nn;
    Gr.n_el = ne;
    Gr.dpn = dpn;
    Gr.npe = npe;
    Gr.n_slot = Lay.n_slot_padded;
    Gr.n_tmpl = in.n_tmpl;
    Gr.n_cls = Lay.n_cls;
    Gr.uniform = in.node_dof == nullptr;
    // alpha and beta per group node when each node's DOFs share them
    {
      bool per_node = true;
      for (int64_t j = 0; j < nn && per_node; ++j)
        for (int c = 1; c < dpn; ++c)
          if (h->alpha[dof0[j] + c] != h->alpha[dof0[j]] ||
              (h->damped && h->beta_host[dof0[j] + c] != h->beta_host[dof0[j]])) {
            per_node = false;
            break;
          }
      if (per_node) {
        std::vector<double> an(nn), bn(h->damped ? nn : 0);
        for (int64_t j = 0; j < nn; ++j) {
          an[j] = h->alpha[dof0[j]];
          if (h->damped) bn[j] = h->beta_host[dof0[j]];
        }
        Gr.coef_node = 1;
        if (h->prec == BFE_PREC_DOUBLE) {
          Gr.alpha_n = upload(an);
          if (h->damped) Gr.beta_n = upload(bn);
        } else {
          Gr.alpha_n = upload(cast_vec<float>(an));
          if (h->damped) Gr.beta_n = upload(cast_vec<float>(bn));
        }
      }
    }
    Gr.pattern = match_pattern(Lay);
    if (Gr.pattern != 0 && !has_pattern_kernel(Gr.pattern, dpn)) Gr.pattern = 0;
    // element records hold each vertex's first DOF
    std::vector<int32_t> nodes4(Lay.nodes4.size());
    for (size_t i = 0; i < nodes4.size(); ++i) nodes4[i] = dof0[Lay.nodes4[i]];
    Gr.adj = upload(Lay.adj);
    Gr.nodes = upload(nodes4);
    if (!Gr.uniform) Gr.dof0 = upload(dof0);
    std::vector<double> stencil;
    if (Gr.pattern != 0) {
      const int32_t nc = Lay.n_cls;
      // first-DOF offsets of the neighbour classes, from a node with every class present
      Gr.doff.assign(nc, 0);
      int64_t ref = -1;
      for (int64_t v = 0; v < nn && ref < 0; ++v) {
        bool all = true;
        for (int32_t j = 0; j < nc && all; ++j) all = Lay.nbr[static_cast<size_t>(j) * nn + v] >= 0;
        if (all) ref = v;
      }
      if (ref < 0) throw std::invalid_argument("element group: no interior node");
      std::vector<int64_t> noff(nc, 0);  // node-index offsets of the classes
      for (int32_t j = 0; j < nc; ++j) {
        const int32_t w = Lay.nbr[static_cast<size_t>(j) * nn + ref];
        Gr.doff[j] = static_cast<int64_t>(dof0[w]) - dof0[ref];
        noff[j] = static_cast<int64_t>(w) - ref;
      }
      // neighbour table as first DOFs; a node is irregular (bit 31) when a
      // neighbour is not at its class offset, in DOFs or in node numbers
      std::vector<int32_t> nbrdof(Lay.nbr.size());
      std::vector<uint32_t> pm(Lay.present);
      for (int64_t v = 0; v < nn; ++v) {
        bool irregular = false;
        for (int32_t j = 0; j < nc; ++j) {
          const int32_t w = Lay.nbr[static_cast<size_t>(j) * nn + v];
          nbrdof[static_cast<size_t>(j) * nn + v] = w < 0 ? -1 : dof0[w];
          if (w >= 0 && (dof0[w] != dof0[v] + Gr.doff[j] || w != v + noff[j])) irregular = true;
        }
        pm[v] = (pm[v] & 0x3fffffffu) | (irregular ? 0x80000000u : 0u);
        // bit 30: every DOF of the node undamped (beta = 2 exactly)
        bool undamped = true;
        if (h->damped)
          for (int c = 0; c < dpn; ++c) undamped = undamped && h->beta_host[dof0[v] + c] == 2.0;
        if (undamped) pm[v] |= 0x40000000u;
      }
      // slots whose element scale is not exactly 1 (absorbing layer)
      std::vector<uint32_t> sm(nn, 0u);
      for (int32_t k = 0; k < Lay.n_slot; ++k)
        for (int64_t v = 0; v < nn; ++v) {
          const int32_t e = Lay.adj[static_cast<size_t>(k) * nn + v];
          if (e >= 0 && in.el_scale[e] != 1.0) sm[v] |= 1u << k;
        }
      Gr.nbr = upload(nbrdof);
      Gr.pmask = upload(pm);
      Gr.smask = upload(sm);
      // most occupied slots scaled (e.g. every triangle, whose scale carries
      // J_rep / J_e): every occupied slot reads its scale from an anchor table
      int64_t occupied = 0, scaled = 0;
      for (int64_t v = 0; v < nn; ++v) {
        occupied += __builtin_popcount(pm[v] & 0x3fffffffu);
        scaled += __builtin_popcount(sm[v]);
      }
      Gr.smask_all = 2 * scaled > occupied;
      // anchor table: element e's scale at (row of its template, node of its
      // canonical first vertex), if every slot view of e agrees on that place
      // and no two elements share one
      {
        const int32_t acls_self = Lay.self_cls;
        std::vector<int32_t> trow(in.n_tmpl, -1);
        int32_t n_rows = 0;
        std::vector<int32_t> srow(Lay.n_slot, -1);
        std::vector<int64_t> el_pos(static_cast<size_t>(ne), -1);
        bool ok = Gr.smask_all && pattern_has_anchor(Gr.pattern);
        for (int32_t k = 0; k < Lay.n_slot && ok; ++k) {
          const int32_t ac = Lay.cls[static_cast<size_t>(k) * npe];
          for (int64_t v = 0; v < nn && ok; ++v) {
            const int32_t e = Lay.adj[static_cast<size_t>(k) * nn + v];
            if (e < 0) continue;
            const int32_t tau = in.el_tmpl[e];
            if (trow[tau] < 0) trow[tau] = n_rows++;
            if (srow[k] < 0) srow[k] = trow[tau];
            const int64_t a = ac == acls_self ? v : Lay.nbr[static_cast<size_t>(ac) * nn + v];
            if (srow[k] != trow[tau] || a < 0) {
              ok = false;
              break;
            }
            const int64_t pos = static_cast<int64_t>(trow[tau]) * nn + a;
            if (el_pos[e] < 0) el_pos[e] = pos;
            else if (el_pos[e] != pos) ok = false;
          }
        }
        std::vector<double> anc;
        // (the kernels index the table with 32-bit offsets)
        ok = ok && static_cast<int64_t>(n_rows) * nn < (int64_t(1) << 31);
        if (ok) {
          anc.assign(static_cast<size_t>(n_rows) * nn, 0.0);
          std::vector<int32_t> taken(static_cast<size_t>(n_rows) * nn, -1);
          for (int64_t e = 0; e < ne && ok; ++e) {
            if (el_pos[e] < 0) continue;
            if (taken[el_pos[e]] >= 0) ok = false;
            taken[el_pos[e]] = static_cast<int32_t>(e);
            anc[el_pos[e]] = in.el_scale[e];
          }
        }
        if (ok) {
          Gr.anc_noff.assign(Lay.n_slot, 0);
          Gr.anc_row.assign(Lay.n_slot, 0);
          for (int32_t k = 0; k < Lay.n_slot; ++k) {
            const int32_t ac = Lay.cls[static_cast<size_t>(k) * npe];
            Gr.anc_noff[k] = static_cast<int32_t>(ac == acls_self ? 0 : noff[ac]);
            Gr.anc_row[k] = std::max(srow[k], 0);
          }
          if (h->prec == BFE_PREC_DOUBLE) Gr.anchor = upload(anc);
          else Gr.anchor = upload(cast_vec<float>(anc));
        }
      }
      // structurally sparse slot rows: every entry outside the pattern's
      // nonzero mask is 0 or cancellation residue (<= 1e-14 of the largest)
      if (Gr.pattern == PatKuhn24::ID && dpn == 3) {
        bool ok = Lay.n_slot == PatKuhn24::NS;
        const size_t m = static_cast<size_t>(PatKuhn24::NS) * dpn * KE;
        double big = 0.0;
        for (size_t i = 0; ok && i < m; ++i) big = std::max(big, std::fabs(Lay.srows_c[i]));
        for (int32_t k = 0; k < Lay.n_slot && ok; ++k)
          for (int c = 0; c < 3 && ok; ++c)
            for (int col = 0; col < KE && ok; ++col)
              ok = PatKuhn24::nzbit(k, c, col) ||
                   std::fabs(Lay.srows_c[(static_cast<size_t>(k) * dpn + c) * KE + col]) <=
                       1e-14 * big;
        Gr.sparse = ok;
      }
      // summed stencil: B_j = sum over (slot, vertex) of class j
      stencil.assign(static_cast<size_t>(nc) * dpn * dpn, 0.0);
      for (int32_t k = 0; k < Lay.n_slot; ++k)
        for (int b = 0; b < npe; ++b) {
          const int32_t j = Lay.cls[k * npe + b];
          for (int c = 0; c < dpn; ++c)
            for (int d = 0; d < dpn; ++d)
              stencil[(static_cast<size_t>(j) * dpn + c) * dpn + d] +=
                  Lay.srows_c[(static_cast<size_t>(k) * dpn + c) * KE + b * dpn + d];
        }
    }
    std::vector<double> scale(in.el_scale, in.el_scale + ne);
    if (h->prec == BFE_PREC_DOUBLE) {
      Gr.scale = upload(scale);
      Gr.srows = upload(Lay.srows);
      if (Gr.pattern != 0) {
        Gr.srows_c = upload(Lay.srows_c);
        Gr.stencil = upload(stencil);
      }
    } else {
      Gr.scale = upload(cast_vec<float>(scale));
      Gr.srows = upload(cast_vec<float>(Lay.srows));
      if (Gr.pattern != 0) {
        Gr.srows_c = upload(cast_vec<float>(Lay.srows_c));
        Gr.stencil = upload(cast_vec<float>(stencil));
      }
    }
    E->groups.push_back(std::move(Gr));
  }
  for (int64_t r = 0; r < n; ++r)
    if (E->row_group[r] < 0)
      throw std::invalid_argument("element groups do not cover every DOF");
  // interface rows of each group (group-local rows), grouped per workgroup
  for (int32_t gi = 0; gi < n_groups; ++gi) {
    EbeGroup& Gr = E->groups[gi];
    std::vector<int32_t> rows, ks;
    for (size_t k = 0; k < h->iface_host.size(); ++k) {
      const int32_t r = h->iface_host[k];
      if (E->row_group[r] == gi) {
        rows.push_back(E->row_local[r]);
        ks.push_back(static_cast<int32_t>(k));
      }
    }
    Grouped g = group_by_block(ebe_blocks(Gr), static_cast<int64_t>(kBlock) * Gr.dpn, rows, &ks,
                               nullptr);
    Gr.iptr = upload(g.ptr);
    Gr.irow = upload(g.row);
    Gr.ik = upload(g.a);
  }
  // per-DOF alpha and beta for groups whose nodes' DOFs do not share them
  bool per_dof = false, any_pattern = false;
  for (const auto& g : E->groups) {
    per_dof = per_dof || !g.coef_node;
    any_pattern = any_pattern || g.pattern != 0;
  }
  double err;
  if (h->prec == BFE_PREC_DOUBLE) {
    if (per_dof) {
      E->alpha = upload(h->alpha);
      if (h->damped) E->beta = upload(h->beta_host);
    }
    err = compare_ebe_with_csr<double>(*h, *E, true);
    if (any_pattern) err = std::max(err, compare_ebe_with_csr<double>(*h, *E, false));
  } else {
    if (per_dof) {
      E->alpha = upload(cast_vec<float>(h->alpha));
      if (h->damped) E->beta = upload(cast_vec<float>(h->beta_host));
    }
    err = compare_ebe_with_csr<float>(*h, *E, true);
    if (any_pattern) err = std::max(err, compare_ebe_with_csr<float>(*h, *E, false));
  }
  E->check_err = err;
  const double tol = check_tol > 0.0 ? check_tol : (h->prec == BFE_PREC_DOUBLE ? 1e-9 : 1e-3);
  if (!(err <= tol))
    throw std::invalid_argument("element operator does not reproduce K (relative step "
                                "difference " + std::to_string(err) + ")");
  h->ebe = std::move(E);
  h->use_ebe = 1;
}

}  // namespace

extern "C" {

bfe_status bfe_create(int64_t n, const int64_t* k_rowptr, const int32_t* k_col,
                      const double* k_val, const double* m, const int64_t* c_rowptr,
                      const int32_t* c_col, const double* c_val, double dt, int32_t mode,
                      int32_t precision, int32_t device, bfe_handle* out) {
  return guarded([&] {
    if (out == nullptr) throw std::invalid_argument("output handle pointer is NULL");
    *out = nullptr;
    if (n <= 0 || n > INT32_MAX) throw std::invalid_argument("n must be in [1, 2^31)");
    if (k_rowptr == nullptr) throw std::invalid_argument("K row pointer is NULL");
    if (precision != BFE_PREC_DOUBLE && precision != BFE_PREC_SINGLE)
      throw std::invalid_argument("unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      throw hip_failure("no HIP device available");
    }
    if (device < 0 || device >= ndev) throw std::invalid_argument("device ordinal out of range");
    // K in place when it is already normal (sorted, no duplicates or exact
    // zeros: what the assemblers produce), else a normalised copy
    bfe::Csr Kc;
    bfe::CsrRef K{n, k_rowptr, k_col, k_val};
    if (!bfe::csr_is_normal(n, k_rowptr, k_col, k_val)) {
      Kc = bfe::normalise_csr(n, k_rowptr, k_col, k_val, "K");
      K = bfe::ref(Kc);
    }
    bfe::Csr C = bfe::normalise_csr(n, c_rowptr, c_col, c_val, "C");
    bfe::Plan P = bfe::build_plan(n, K, m, C, dt, mode, /*host_sell=*/false);

    auto H = std::make_unique<bfe_plan_s>();
    H->device = device;
    BFE_HIP(hipSetDevice(device));
    BFE_HIP(hipStreamCreateWithFlags(&H->stream, hipStreamNonBlocking));
    BFE_HIP(hipEventCreate(&H->ev0));
    BFE_HIP(hipEventCreate(&H->ev1));
    H->n = n;
    H->mode = mode;
    H->prec = precision;
    H->dt = dt;
    H->damped = P.damped;
    H->nnz_k = P.nnz_k;
    H->nI = static_cast<int32_t>(P.iface.size());
    H->nnz_coff = P.coff.nnz();
    H->nnz_xinv = P.has_xinv ? P.xinv.nnz() : 0;
    H->has_xinv = P.has_xinv;
    H->sq = P.sq;
    H->sr = P.sr;
    H->alpha = P.alpha;
    H->iface_host = P.iface;
    H->m_host.assign(m, m + n);
    if (P.damped) H->beta_host = P.beta;
    if (precision == BFE_PREC_DOUBLE) {
      build_operators<double>(*H, K, P.alpha);
      upload_plan<double>(*H, P);
    } else {
      build_operators<float>(*H, K, P.alpha);
      upload_plan<float>(*H, P);
    }
    *out = H.release();
  });
}

bfe_status bfe_run(bfe_handle h, int32_t n_t, const int32_t* frc_idx, int32_t p,
                   const double* frcs, int32_t frc_rows, const int32_t* hist_idx, int32_t q,
                   double* hist_out, int32_t field_every, int32_t field_is_vel,
                   double* field_out, int32_t* ti_start_out) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    if (h->prec == BFE_PREC_DOUBLE)
      run_single<double>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out,
                         field_every, field_is_vel, field_out, ti_start_out, nullptr, 0);
    else
      run_single<float>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out,
                        field_every, field_is_vel, field_out, ti_start_out, nullptr, 0);
  });
}

bfe_status bfe_set_field_elements(bfe_handle h, int64_t n_nodes, int32_t max_df,
                                  const int32_t* node_dof, int64_t n_el, int32_t npe,
                                  const int32_t* els) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    if (n_nodes < 1 || max_df < 1 || n_el < 1 || npe < 1 || npe > 8)
      throw std::invalid_argument("field elements: bad sizes");
    if (node_dof == nullptr || els == nullptr) throw std::invalid_argument("NULL field tables");
    for (int64_t i = 0; i < n_nodes * max_df; ++i)
      if (node_dof[i] < -1 || node_dof[i] >= h->n)
        throw std::invalid_argument("field elements: DOF index out of range");
    for (int64_t i = 0; i < n_el * npe; ++i)
      if (els[i] < -1 || els[i] >= n_nodes)
        throw std::invalid_argument("field elements: node index out of range");
    BFE_HIP(hipSetDevice(h->device));
    auto F = std::make_shared<FieldElements>();
    F->n_el = n_el;
    F->n_nodes = n_nodes;
    F->npe = npe;
    F->max_df = max_df;
    F->els = upload(std::vector<int32_t>(els, els + n_el * npe));
    F->node_dof = upload(std::vector<int32_t>(node_dof, node_dof + n_nodes * max_df));
    if (h->prec == BFE_PREC_DOUBLE) F->m = upload(h->m_host);
    else F->m = upload(cast_vec<float>(h->m_host));
    h->field_elems = F;
  });
}

bfe_status bfe_run_field_ke(bfe_handle h, int32_t n_t, const int32_t* frc_idx, int32_t p,
                            const double* frcs, int32_t frc_rows, const int32_t* hist_idx,
                            int32_t q, double* hist_out, int32_t field_every, double* el_out,
                            int32_t* ti_start_out) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    if (field_every <= 0) throw std::invalid_argument("field_every must be positive");
    if (el_out == nullptr) throw std::invalid_argument("element field output is NULL");
    if (h->prec == BFE_PREC_DOUBLE)
      run_single<double>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out, field_every,
                         1, nullptr, ti_start_out, nullptr, 0, el_out);
    else
      run_single<float>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out, field_every,
                        1, nullptr, ti_start_out, nullptr, 0, el_out);
  });
}

bfe_status bfe_set_field_geometry(bfe_handle h, int32_t dim, const double* nds) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    auto F = std::static_pointer_cast<FieldElements>(h->field_elems);
    if (!F) throw std::invalid_argument("field elements not set");
    if (dim < 1 || dim > 3) throw std::invalid_argument("dim must be 1, 2 or 3");
    if (dim > F->max_df) throw std::invalid_argument("dim exceeds the DOFs per node");
    if (nds == nullptr) throw std::invalid_argument("NULL node coordinates");
    for (int64_t i = 0; i < F->n_nodes * dim; ++i)
      if (!std::isfinite(nds[i])) throw std::invalid_argument("non-finite node coordinate");
    BFE_HIP(hipSetDevice(h->device));
    F->nds = upload(std::vector<double>(nds, nds + F->n_nodes * dim));
    F->dim = dim;
  });
}

bfe_status bfe_run_field(bfe_handle h, int32_t n_t, const int32_t* frc_idx, int32_t p,
                         const double* frcs, int32_t frc_rows, const int32_t* hist_idx, int32_t q,
                         double* hist_out, int32_t field_every, int32_t field_type,
                         int32_t field_is_vel, int32_t comp, double* el_out,
                         int32_t* ti_start_out) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    if (field_every <= 0) throw std::invalid_argument("field_every must be positive");
    if (el_out == nullptr) throw std::invalid_argument("element field output is NULL");
    auto F = std::static_pointer_cast<FieldElements>(h->field_elems);
    if (!F) throw std::invalid_argument("field elements not set");
    if (field_type < BFE_FIELD_KE || field_type > BFE_FIELD_CURL)
      throw std::invalid_argument("unknown field type");
    if (field_type == BFE_FIELD_MEAN && (comp < 0 || comp >= F->max_df))
      throw std::invalid_argument("field component out of range");
    if (field_type == BFE_FIELD_DIV || field_type == BFE_FIELD_CURL) {
      if (F->dim == 0) throw std::invalid_argument("field geometry not set");
      if (field_type == BFE_FIELD_CURL && F->dim != 2)  // fn_get_field_output.m:47-49
        throw std::invalid_argument("Cannot output curl for 3D models");
    }
    const int32_t vel = field_is_vel ? 1 : 0;
    if (h->prec == BFE_PREC_DOUBLE)
      run_single<double>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out, field_every,
                         vel, nullptr, ti_start_out, nullptr, 0, el_out, field_type, comp);
    else
      run_single<float>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out, field_every,
                        vel, nullptr, ti_start_out, nullptr, 0, el_out, field_type, comp);
  });
}

bfe_status bfe_run_checkpointed(bfe_handle h, int32_t slot, int32_t n_t, const int32_t* frc_idx,
                                int32_t p, const double* frcs, int32_t frc_rows,
                                const int32_t* hist_idx, int32_t q, double* hist_out,
                                int32_t ckpt_every, int32_t* ti_start_out) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    if (ckpt_every < 1) throw std::invalid_argument("checkpoint interval must be >= 1");
    if (slot < 0) throw std::invalid_argument("slot must be >= 0");
    auto rec = std::make_shared<AdjointRecord>();
    h->adj_records.erase(slot);
    if (h->prec == BFE_PREC_DOUBLE)
      run_single<double>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out, 0, 1,
                         nullptr, ti_start_out, rec.get(), ckpt_every);
    else
      run_single<float>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out, 0, 1,
                        nullptr, ti_start_out, rec.get(), ckpt_every);
    h->adj_records[slot] = rec;
  });
}

bfe_status bfe_run_checkpointed_batched(bfe_handle h, int32_t slot, int32_t n_t, int32_t n_src,
                                        const int32_t* frc_idx, const int32_t* frc_src, int32_t p,
                                        const double* frcs, int32_t frc_rows,
                                        const int32_t* hist_idx, int32_t q, double* hist_out,
                                        int32_t ckpt_every, int32_t* ti_start_out) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    if (ckpt_every < 1) throw std::invalid_argument("checkpoint interval must be >= 1");
    if (slot < 0) throw std::invalid_argument("slot must be >= 0");
    if (n_src < 1 || n_src > 32)
      throw std::invalid_argument("a batched checkpointed run takes 1 to 32 sources");
    auto rec = std::make_shared<AdjointRecord>();
    h->adj_records.erase(slot);
    if (h->prec == BFE_PREC_DOUBLE)
      run_batched<double>(*h, n_t, n_src, frc_idx, frc_src, p, frcs, frc_rows, hist_idx, q,
                          hist_out, ti_start_out, rec.get(), ckpt_every);
    else
      run_batched<float>(*h, n_t, n_src, frc_idx, frc_src, p, frcs, frc_rows, hist_idx, q,
                         hist_out, ti_start_out, rec.get(), ckpt_every);
    h->adj_records[slot] = rec;
  });
}

bfe_status bfe_set_gradient_elements(bfe_handle h, int32_t n_el, int32_t kmax,
                                     const int32_t* el_dofs, const double* el_K,
                                     const double* el_m, const double* el_mu) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    if (n_el < 1 || kmax < 1) throw std::invalid_argument("n_el and kmax must be positive");
    if (!el_dofs || !el_K || !el_m || !el_mu) throw std::invalid_argument("NULL element array");
    const size_t ne = static_cast<size_t>(n_el);
    for (size_t i = 0; i < ne * kmax; ++i)
      if (el_dofs[i] < -1 || el_dofs[i] >= h->n)
        throw std::invalid_argument("element DOF index out of range");
    BFE_HIP(hipSetDevice(h->device));
    auto G = std::make_shared<GradElements>();
    G->n_el = n_el;
    G->kmax = kmax;
    // mass-term weights of dR/drho_e per solver mode (see run_adjoint)
    std::vector<double> w(3 * ne);
    const double dt = h->dt;
    for (size_t e = 0; e < ne; ++e) {
      const double mu = el_mu[e];
      if (h->mode == BFE_MODE_VEL_AT_CURRENT) {
        w[3 * e] = 1.0 + 0.5 * dt * mu; w[3 * e + 1] = -2.0; w[3 * e + 2] = 1.0 - 0.5 * dt * mu;
      } else {
        w[3 * e] = 1.0; w[3 * e + 1] = -(2.0 - dt * mu); w[3 * e + 2] = 1.0 - dt * mu;
      }
    }
    std::vector<int32_t> d(el_dofs, el_dofs + ne * kmax);
    std::vector<double> K(el_K, el_K + ne * kmax * kmax), m(el_m, el_m + ne * kmax);
    G->dofs = upload(d);
    // node-pair form for the batched adjoint: nodes of 2 DOFs, else of 1
    if (!pair_layout(d, static_cast<int64_t>(ne), kmax, h->n, 2, 4, *G))
      (void)pair_layout(d, static_cast<int64_t>(ne), kmax, h->n, 1, 6, *G);
    if (h->prec == BFE_PREC_DOUBLE) {
      G->K = upload(K); G->m = upload(m); G->w = upload(w);
    } else {
      G->K = upload(cast_vec<float>(K)); G->m = upload(cast_vec<float>(m));
      G->w = upload(cast_vec<float>(w));
    }
    h->grad_elems = G;
  });
}

bfe_status bfe_adjoint(bfe_handle h, int32_t slot, const double* hist_grad, double* g_stiffness,
                       double* g_density) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    auto it = h->adj_records.find(slot);
    auto rec = it == h->adj_records.end() ? nullptr
                                          : std::static_pointer_cast<AdjointRecord>(it->second);
    auto G = std::static_pointer_cast<GradElements>(h->grad_elems);
    if (!rec || !rec->valid)
      throw std::invalid_argument("no checkpointed forward run to differentiate");
    if (!G) throw std::invalid_argument("gradient elements not set");
    if (!g_stiffness || !g_density) throw std::invalid_argument("NULL gradient output");
    if (rec->q > 0 && hist_grad == nullptr) throw std::invalid_argument("NULL history gradient");
    const bool batched = rec->batched;
    if (h->prec == BFE_PREC_DOUBLE) {
      if (batched) run_adjoint_batched<double>(*h, *rec, *G, hist_grad, g_stiffness, g_density);
      else run_adjoint<double>(*h, *rec, *G, hist_grad, g_stiffness, g_density);
    } else {
      if (batched) run_adjoint_batched<float>(*h, *rec, *G, hist_grad, g_stiffness, g_density);
      else run_adjoint<float>(*h, *rec, *G, hist_grad, g_stiffness, g_density);
    }
    h->adj_records.erase(slot);  // the forward states were consumed
  });
}

bfe_status bfe_run_batched(bfe_handle h, int32_t n_t, int32_t n_src, const int32_t* frc_idx,
                           const int32_t* frc_src, int32_t p, const double* frcs,
                           int32_t frc_rows, const int32_t* hist_idx, int32_t q,
                           double* hist_out, int32_t* ti_start_out) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    if (h->prec == BFE_PREC_DOUBLE)
      run_batched<double>(*h, n_t, n_src, frc_idx, frc_src, p, frcs, frc_rows, hist_idx, q,
                          hist_out, ti_start_out);
    else
      run_batched<float>(*h, n_t, n_src, frc_idx, frc_src, p, frcs, frc_rows, hist_idx, q,
                         hist_out, ti_start_out);
  });
}

bfe_status bfe_set_element_groups(bfe_handle h, int32_t n_groups, const bfe_element_group* groups,
                                  double check_tol) {
  return guarded([&] { set_element_groups(h, n_groups, groups, check_tol); });
}

bfe_status bfe_set_element_operator(bfe_handle h, int64_t n_nodes, int32_t dpn, int64_t n_el,
                                    int32_t npe, const int32_t* el_nodes, const int32_t* el_tmpl,
                                    const double* el_scale, int32_t n_tmpl, const double* tmpl,
                                    double check_tol) {
  return guarded([&] {
    bfe_element_group g{};
    g.n_nodes = n_nodes;
    g.dpn = dpn;
    g.npe = npe;
    g.n_el = n_el;
    g.el_nodes = el_nodes;
    g.el_tmpl = el_tmpl;
    g.el_scale = el_scale;
    g.n_tmpl = n_tmpl;
    g.tmpl = tmpl;
    g.node_dof = nullptr;  // node j's DOFs are j*dpn + c
    set_element_groups(h, 1, &g, check_tol);
  });
}

bfe_status bfe_set_tuning(bfe_handle h, int32_t key, int32_t value) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    switch (key) {
      case BFE_TUNE_MULTI_ROWS_PER_BLOCK:
        if (value != 1 && value != 2 && value != 4 && value != 8)
          throw std::invalid_argument("rows per workgroup must be 1, 2, 4 or 8");
        h->multi_rows = value;
        break;
      case BFE_TUNE_MULTI_ROWS_PER_WAVE:
        if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16)
          throw std::invalid_argument("rows per wavefront must be 1, 2, 4, 8 or 16");
        h->multi_rpw = value;
        break;
      case BFE_TUNE_MULTI_GROUP:
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
          throw std::invalid_argument("rows per gather group must be 0 (automatic), 1, 2, 4 or 8");
        h->multi_group = value;
        break;
      case BFE_TUNE_XCD_REMAP:
        h->xcd_remap = value ? 1 : 0;
        break;
      case BFE_TUNE_EBE_GROUP:
        if (value != 2 && value != 4 && value != 8)
          throw std::invalid_argument("element-operator load group must be 2, 4 or 8");
        h->ebe_group = value;
        break;
      case BFE_TUNE_GRAPH:
        h->use_graph = value ? 1 : 0;
        break;
      case BFE_TUNE_PERSISTENT:
        h->use_persistent = value ? 1 : 0;
        break;
      case BFE_TUNE_MULTI_OPERATOR:
        if (value != BFE_OPERATOR_CSR && value != BFE_OPERATOR_EBE && value != BFE_OPERATOR_NODE)
          throw std::invalid_argument(
              "operator must be BFE_OPERATOR_CSR, BFE_OPERATOR_EBE or BFE_OPERATOR_NODE");
        h->multi_ebe = value;
        break;
      case BFE_TUNE_MULTI_NODES_PER_WAVE:
        if (value < 1 || value > 1024)
          throw std::invalid_argument("nodes per wavefront must be in [1, 1024]");
        h->multi_npw = value;
        break;
      case BFE_TUNE_MULTI_SLIDE:
        h->multi_pipeline = value ? 1 : 0;
        break;
      case BFE_TUNE_ADJOINT_PAIRS:
        h->grad_pairs = value ? 1 : 0;
        break;
      case BFE_TUNE_PERSISTENT_GRID:
        if (value < 0) throw std::invalid_argument("persistent grid cap must be >= 0");
        h->persistent_grid = value;
        break;
      case BFE_TUNE_EBE_KERNEL:
        h->ebe_kernel = value ? 1 : 0;
        break;
      case BFE_TUNE_OPERATOR:
        if (value != BFE_OPERATOR_CSR && value != BFE_OPERATOR_EBE)
          throw std::invalid_argument("operator must be BFE_OPERATOR_CSR or BFE_OPERATOR_EBE");
        if (value == BFE_OPERATOR_EBE && !h->ebe)
          throw std::invalid_argument("no element operator attached");
        h->use_ebe = value == BFE_OPERATOR_EBE ? 1 : 0;
        break;
      default:
        throw std::invalid_argument("unknown tuning key");
    }
  });
}

bfe_status bfe_set_kernel_timing(bfe_handle h, int32_t every) {
  return guarded([&] {
    if (h == nullptr) throw std::invalid_argument("NULL handle");
    h->timing_every = every > 0 ? every : 0;
  });
}

bfe_status bfe_info_get(bfe_handle h, bfe_info* info) {
  return guarded([&] {
    if (h == nullptr || info == nullptr) throw std::invalid_argument("NULL argument");
    bfe_info I{};
    I.n = h->n;
    I.nnz_k = h->nnz_k;
    I.nnz_sell = h->nnz_sell;
    I.n_interface = h->nI;
    I.nnz_interface = h->nnz_coff;
    I.nnz_xinv = h->nnz_xinv;
    I.damped = h->damped ? 1 : 0;
    I.mode = h->mode;
    I.precision = h->prec;
    I.device = h->device;
    I.device_bytes = h->device_bytes();
    I.operator_kind = (h->ebe && h->use_ebe) ? BFE_OPERATOR_EBE : BFE_OPERATOR_CSR;
    if (h->ebe) {
      // totals over the groups; pattern and classes of the first group
      const EbeOp& E = *h->ebe;
      for (const auto& g : E.groups) {
        I.ebe_slots += g.n_slot;
        I.ebe_templates += g.n_tmpl;
      }
      I.ebe_elements = E.n_el();
      I.ebe_check = E.check_err;
      I.ebe_pattern = E.groups[0].pattern;
      I.ebe_classes = E.groups[0].n_cls;
      I.ebe_groups = static_cast<int32_t>(E.groups.size());
      for (const auto& g : E.groups) I.ebe_sparse += g.sparse;
    }
    I.multi_group = h->last_group;
    I.multi_operator = h->last_multi_op;
    I.grad_pairs = h->grad_elems ? std::static_pointer_cast<GradElements>(h->grad_elems)->pair_gd : 0;
    *info = I;
  });
}

bfe_status bfe_last_run_stats(bfe_handle h, bfe_run_stats* stats) {
  return guarded([&] {
    if (h == nullptr || stats == nullptr) throw std::invalid_argument("NULL argument");
    *stats = h->stats;
  });
}

bfe_status bfe_destroy(bfe_handle h) {
  return guarded([&] { delete h; });
}

const char* bfe_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

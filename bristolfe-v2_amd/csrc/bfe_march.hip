// MI355X (gfx950) kernels and C ABI of the explicit time-march.
//
// One time step of code/internal/fn_explicit_dynamic_solver_v6.m:180-212
//   f(idx) = F(:, ti);  u = A f + B u1 + D u2;  hist(:, ti) = u(h_idx);
//   [fld(:, fi) = (u - u1)/dt];  u2 = u1;  u1 = u
// becomes one fused, HBM-streaming kernel over all rows (k_step) plus, only
// when the model has fluid-solid interface rows, one single-workgroup kernel
// (k_iface) that applies the interface coupling and X_II^-1. Monitor gathers
// and field frames of step t-1 are written by the step-t kernel from the
// u1/u2 values it already reads, so they cost no extra launch or pass.
//
// Operator layout (host set-up in bfe_plan.cpp): SELL-64 of alpha-scaled K.
// A wavefront owns one 64-row slice; lane l streams row 64*s+l, so the value
// and column streams are read fully coalesced (64 consecutive entries per
// load) and no cross-lane reduction is needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bfe.h"
#include "bfe_plan.hpp"

namespace {

constexpr int kBlock = 256;    // rows per workgroup in k_step (4 slices)
constexpr int kIfaceThreads = 1024;

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

// ---------------------------------------------------------------------------
// Single right-hand side step.
//   u[r] = beta[r] u1[r] + gamma[r] u2[r] - sum_k val[k] u1[col[k]]  (+ forcing)
// DAMPED = false: beta = 2, gamma = -1 for every row (no arrays read).
template <typename T, bool DAMPED>
__global__ __launch_bounds__(kBlock) void k_step(
    int64_t n, int64_t nblk, const int64_t* __restrict__ off,
    const int32_t* __restrict__ col, const T* __restrict__ val,
    const T* __restrict__ beta, const T* __restrict__ gamma,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u,
    const int32_t* __restrict__ fptr, const int32_t* __restrict__ f_row,
    const T* __restrict__ f_alpha, const int32_t* __restrict__ f_frow,
    const T* __restrict__ F, int32_t F_rows, int32_t t,
    const int32_t* __restrict__ mptr, const int32_t* __restrict__ m_row,
    const int32_t* __restrict__ m_k, T* __restrict__ hist_prev,
    T* __restrict__ fld_prev, int32_t fld_is_vel, T dt) {
  const int64_t lb = xcd_block(blockIdx.x, nblk);
  const int64_t row = lb * kBlock + threadIdx.x;
  if (row >= n) return;
  const int64_t sl = row >> 6;
  const int lane = static_cast<int>(row & 63);
  const int64_t e = off[sl + 1];
  int64_t p = off[sl] + lane;
  T acc0 = T(0), acc1 = T(0);
  for (; p + 64 < e; p += 128) {
    acc0 += val[p] * u1[col[p]];
    acc1 += val[p + 64] * u1[col[p + 64]];
  }
  if (p < e) acc0 += val[p] * u1[col[p]];
  const T a = u1[row], b = u2[row];
  T un;
  if constexpr (DAMPED) un = beta[row] * a + gamma[row] * b - (acc0 + acc1);
  else un = (T(2) * a - b) - (acc0 + acc1);
  for (int32_t i = fptr[lb], ie = fptr[lb + 1]; i < ie; ++i)
    if (f_row[i] == row) un += f_alpha[i] * F[static_cast<int64_t>(t) * F_rows + f_frow[i]];
  u[row] = un;
  if (fld_prev != nullptr) fld_prev[row] = fld_is_vel ? (a - b) / dt : a;
  if (hist_prev != nullptr)
    for (int32_t i = mptr[lb], ie = mptr[lb + 1]; i < ie; ++i)
      if (m_row[i] == row) hist_prev[m_k[i]] = a;
}

// ---------------------------------------------------------------------------
// Multi right-hand side step (S <= 64 sources, u stored row-major n x S).
// One wavefront per row, lane = source: the row's K entries are wave-uniform
// and the u1 gathers read S contiguous values.
template <typename T, bool DAMPED>
__global__ __launch_bounds__(kBlock) void k_step_multi(
    int64_t n, int32_t S, int64_t nblk, const int64_t* __restrict__ off,
    const int32_t* __restrict__ col, const T* __restrict__ val,
    const T* __restrict__ beta, const T* __restrict__ gamma,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u,
    const int32_t* __restrict__ fptr, const int32_t* __restrict__ f_row,
    const int32_t* __restrict__ f_src, const T* __restrict__ f_alpha,
    const int32_t* __restrict__ f_frow, const T* __restrict__ F, int32_t F_rows,
    int32_t t, const int32_t* __restrict__ mptr, const int32_t* __restrict__ m_row,
    const int32_t* __restrict__ m_k, int32_t q, T* __restrict__ hist_prev) {
  const int64_t lb = xcd_block(blockIdx.x, nblk);
  const int64_t row = lb * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n || lane >= S) return;
  const int64_t sl = row >> 6;
  const int64_t e = off[sl + 1];
  T acc = T(0);
  for (int64_t p = off[sl] + (row & 63); p < e; p += 64)
    acc += val[p] * u1[static_cast<int64_t>(col[p]) * S + lane];
  const int64_t me = row * S + lane;
  const T a = u1[me], b = u2[me];
  T un;
  if constexpr (DAMPED) un = beta[row] * a + gamma[row] * b - acc;
  else un = (T(2) * a - b) - acc;
  for (int32_t i = fptr[lb], ie = fptr[lb + 1]; i < ie; ++i)
    if (f_row[i] == row && f_src[i] == lane)
      un += f_alpha[i] * F[static_cast<int64_t>(t) * F_rows + f_frow[i]];
  u[me] = un;
  if (hist_prev != nullptr)
    for (int32_t i = mptr[lb], ie = mptr[lb + 1]; i < ie; ++i)
      if (m_row[i] == row) hist_prev[m_k[i] + static_cast<int64_t>(q) * lane] = a;
}

// ---------------------------------------------------------------------------
// Interface rows (one workgroup). For interface entry k (global row r):
//   y = w_r + sq * sum_j Coff_kj u1_j + sr * sum_j Coff_kj u2_j (+ forcing)
//   u_r = sum_l Xinv_kl y_l        (identity when xptr == nullptr)
// w_r is what k_step wrote to u[r] with the interface-row coefficients.
template <typename T>
__global__ __launch_bounds__(kIfaceThreads) void k_iface(
    int32_t nI, int32_t S, const int32_t* __restrict__ iface,
    const int64_t* __restrict__ cptr, const int32_t* __restrict__ ccol,
    const T* __restrict__ cval, T sq, T sr, const int64_t* __restrict__ xptr,
    const int32_t* __restrict__ xcol, const T* __restrict__ xval,
    const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ u,
    T* __restrict__ y, int32_t nfi, const int32_t* __restrict__ fi_k,
    const int32_t* __restrict__ fi_src, const T* __restrict__ fi_alpha,
    const int32_t* __restrict__ fi_frow, const T* __restrict__ F, int32_t F_rows,
    int32_t t) {
  const int64_t tot = static_cast<int64_t>(nI) * S;
  for (int64_t idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int64_t k = idx / S, s = idx - k * S;
    const int64_t r = iface[k];
    T a1 = T(0), a2 = T(0);
    for (int64_t p = cptr[k], pe = cptr[k + 1]; p < pe; ++p) {
      const int64_t c = static_cast<int64_t>(ccol[p]) * S + s;
      a1 += cval[p] * u1[c];
      a2 += cval[p] * u2[c];
    }
    y[idx] = u[r * S + s] + sq * a1 + sr * a2;
  }
  __syncthreads();
  for (int32_t i = threadIdx.x; i < nfi; i += blockDim.x)
    y[static_cast<int64_t>(fi_k[i]) * S + fi_src[i]] +=
        fi_alpha[i] * F[static_cast<int64_t>(t) * F_rows + fi_frow[i]];
  __syncthreads();
  for (int64_t idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int64_t k = idx / S, s = idx - k * S;
    T acc;
    if (xptr != nullptr) {
      acc = T(0);
      for (int64_t p = xptr[k], pe = xptr[k + 1]; p < pe; ++p)
        acc += xval[p] * y[static_cast<int64_t>(xcol[p]) * S + s];
    } else {
      acc = y[idx];
    }
    u[static_cast<int64_t>(iface[k]) * S + s] = acc;
  }
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

template <typename U>
DevBuf upload(const std::vector<U>& v, hipStream_t st) {
  DevBuf b(v.size() * sizeof(U));
  if (!v.empty()) BFE_HIP(hipMemcpyAsync(b.p, v.data(), b.bytes, hipMemcpyHostToDevice, st));
  return b;
}

template <typename T>
std::vector<T> cast_vec(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

}  // namespace

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
  double sq = 0.0, sr = 0.0;
  bool has_xinv = false;
  std::vector<double> alpha;
  std::vector<int32_t> iface_pos;
  DevBuf off, col, val, beta, gamma, iface, cptr, ccol, cval, xptr, xcol, xval;
  bfe_run_stats stats{};
  int32_t timing_every = 0;
  std::vector<hipEvent_t> tev;  // pairs (start, stop) for sampled launches
  ~bfe_plan_s() {
    if (stream) {
      (void)hipSetDevice(device);
      (void)hipStreamSynchronize(stream);
    }
    off.release(); col.release(); val.release(); beta.release(); gamma.release();
    iface.release(); cptr.release(); ccol.release(); cval.release();
    xptr.release(); xcol.release(); xval.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
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
  hipStream_t st = H.stream;
  H.off = upload(P.op.off, st);
  H.col = upload(P.op.col, st);
  H.val = upload(cast_vec<T>(P.op.val), st);
  if (P.damped) {
    H.beta = upload(cast_vec<T>(P.beta), st);
    H.gamma = upload(cast_vec<T>(P.gamma), st);
  }
  if (!P.iface.empty()) {
    H.iface = upload(P.iface, st);
    H.cptr = upload(P.coff.ptr, st);
    H.ccol = upload(P.coff.col, st);
    H.cval = upload(cast_vec<T>(P.coff.val), st);
    if (P.has_xinv) {
      H.xptr = upload(P.xinv.ptr, st);
      H.xcol = upload(P.xinv.col, st);
      H.xval = upload(cast_vec<T>(P.xinv.val), st);
    }
  }
  BFE_HIP(hipStreamSynchronize(st));
}

// Host-side lists consumed by the step kernels.
struct RunLists {
  // forcing on non-interface rows, grouped by logical block
  std::vector<int32_t> fptr, f_row, f_src, f_frow;
  std::vector<double> f_alpha;
  // forcing on interface rows
  std::vector<int32_t> fi_k, fi_src, fi_frow;
  std::vector<double> fi_alpha;
  // monitors grouped by logical block
  std::vector<int32_t> mptr, m_row, m_k;
};

// Forcing entries: (dof, source column, signal row). For one source column
// the last entry for a DOF wins, like MATLAB's f(idx) = v.
RunLists make_lists(const bfe_plan_s& H, int64_t rows_per_block, int64_t nblk,
                    const std::vector<int32_t>& fdof, const std::vector<int32_t>& fsrc,
                    const std::vector<int32_t>& frow, const int32_t* hist_idx, int32_t q) {
  RunLists L;
  // dedupe (dof, src) keeping the last occurrence
  std::vector<std::pair<std::pair<int32_t, int32_t>, int32_t>> key;
  key.reserve(fdof.size());
  for (size_t i = 0; i < fdof.size(); ++i) key.push_back({{fdof[i], fsrc[i]}, static_cast<int32_t>(i)});
  std::stable_sort(key.begin(), key.end(),
                   [](auto& a, auto& b) { return a.first < b.first; });
  std::vector<int32_t> keep;
  for (size_t i = 0; i < key.size(); ++i)
    if (i + 1 == key.size() || key[i + 1].first != key[i].first) keep.push_back(key[i].second);
  std::vector<std::vector<int32_t>> per_block_f;
  std::vector<int64_t> fcount(nblk + 1, 0);
  std::vector<int32_t> blk_entries;
  for (int32_t i : keep) {
    const int32_t d = fdof[i];
    if (H.nI > 0 && H.iface_pos[d] >= 0) {
      L.fi_k.push_back(H.iface_pos[d]);
      L.fi_src.push_back(fsrc[i]);
      L.fi_frow.push_back(frow[i]);
      L.fi_alpha.push_back(H.alpha[d]);
    } else {
      blk_entries.push_back(i);
      fcount[d / rows_per_block + 1]++;
    }
  }
  L.fptr.assign(nblk + 1, 0);
  for (int64_t b = 0; b < nblk; ++b) L.fptr[b + 1] = static_cast<int32_t>(L.fptr[b] + fcount[b + 1]);
  const size_t nf = blk_entries.size();
  L.f_row.resize(nf); L.f_src.resize(nf); L.f_frow.resize(nf); L.f_alpha.resize(nf);
  std::vector<int32_t> fill(L.fptr.begin(), L.fptr.end() - 1);
  for (int32_t i : blk_entries) {
    const int32_t d = fdof[i];
    const int32_t o = fill[d / rows_per_block]++;
    L.f_row[o] = d; L.f_src[o] = fsrc[i]; L.f_frow[o] = frow[i]; L.f_alpha[o] = H.alpha[d];
  }
  // monitors
  L.mptr.assign(nblk + 1, 0);
  std::vector<int64_t> mcount(nblk + 1, 0);
  for (int32_t k = 0; k < q; ++k) mcount[hist_idx[k] / rows_per_block + 1]++;
  for (int64_t b = 0; b < nblk; ++b) L.mptr[b + 1] = static_cast<int32_t>(L.mptr[b] + mcount[b + 1]);
  L.m_row.resize(q); L.m_k.resize(q);
  std::vector<int32_t> mfill(L.mptr.begin(), L.mptr.end() - 1);
  for (int32_t k = 0; k < q; ++k) {
    const int32_t o = mfill[hist_idx[k] / rows_per_block]++;
    L.m_row[o] = hist_idx[k]; L.m_k[o] = k;
  }
  return L;
}

// 0-based first column of F with a non-zero (or NaN) entry; -1 if none.
int64_t first_forced_step(const double* F, int32_t rows, int32_t n_t) {
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

template <typename T>
void run_single(bfe_plan_s& H, int32_t n_t, const int32_t* frc_idx, int32_t p,
                const double* frcs, int32_t frc_rows, const int32_t* hist_idx, int32_t q,
                double* hist_out, int32_t field_every, int32_t field_is_vel,
                double* field_out, int32_t* ti_start_out) {
  const int64_t n = H.n;
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
  const int32_t n_frames = field_every > 0 ? (n_t - 1) / field_every + 1 : 0;
  if (n_frames > 0 && field_out == nullptr) throw std::invalid_argument("field output is NULL");

  if (q > 0) std::memset(hist_out, 0, sizeof(double) * static_cast<size_t>(q) * n_t);
  if (n_frames > 0) std::memset(field_out, 0, sizeof(double) * static_cast<size_t>(n) * n_frames);
  H.stats = bfe_run_stats{};
  H.stats.columns = 1;
  const int64_t t0 = p > 0 ? first_forced_step(frcs, frc_rows, n_t) : -1;
  if (ti_start_out) *ti_start_out = t0 < 0 ? n_t + 1 : static_cast<int32_t>(t0 + 1);
  if (t0 < 0) return;  // empty loop (v6:163-180): outputs stay zero

  BFE_HIP(hipSetDevice(H.device));
  hipStream_t st = H.stream;
  const int64_t nblk = (n + kBlock - 1) / kBlock;
  std::vector<int32_t> fdof(frc_idx, frc_idx + p), fsrc(p, 0), frow(p);
  for (int32_t i = 0; i < p; ++i) frow[i] = frc_rows == 1 ? 0 : i;
  RunLists L = make_lists(H, kBlock, nblk, fdof, fsrc, frow, hist_idx, q);

  std::vector<double> Fh(frcs, frcs + static_cast<size_t>(p > 0 ? frc_rows : 0) * n_t);
  DevBuf dF = upload(cast_vec<T>(Fh), st);
  DevBuf dfptr = upload(L.fptr, st), dfrow = upload(L.f_row, st),
         dffrow = upload(L.f_frow, st), dfalpha = upload(cast_vec<T>(L.f_alpha), st);
  DevBuf dfik = upload(L.fi_k, st), dfisrc = upload(L.fi_src, st),
         dfifrow = upload(L.fi_frow, st), dfialpha = upload(cast_vec<T>(L.fi_alpha), st);
  DevBuf dmptr = upload(L.mptr, st), dmrow = upload(L.m_row, st), dmk = upload(L.m_k, st);
  std::vector<int32_t> hidx(hist_idx, hist_idx + q);
  DevBuf dhidx = upload(hidx, st);
  DevBuf dhist(sizeof(T) * static_cast<size_t>(q) * n_t);
  DevBuf dfld(sizeof(T) * static_cast<size_t>(n) * n_frames);
  DevBuf dy(sizeof(T) * static_cast<size_t>(H.nI));
  DevBuf ubuf(sizeof(T) * static_cast<size_t>(n) * 3);
  BFE_HIP(hipMemsetAsync(ubuf.p, 0, ubuf.bytes, st));
  if (dhist.bytes) BFE_HIP(hipMemsetAsync(dhist.p, 0, dhist.bytes, st));
  if (dfld.bytes) BFE_HIP(hipMemsetAsync(dfld.p, 0, dfld.bytes, st));
  T* U[3] = {ubuf.as<T>(), ubuf.as<T>() + n, ubuf.as<T>() + 2 * n};
  T *u = U[0], *u1 = U[1], *u2 = U[2];
  T* hist = dhist.as<T>();
  T* fld = dfld.as<T>();
  const T dtT = static_cast<T>(H.dt);
  int64_t launches = 0;
  KernelSampler sampler(H, t0, n_t - t0);

  BFE_HIP(hipEventRecord(H.ev0, st));
  for (int64_t t = t0; t < n_t; ++t) {
    const int64_t prev = t - 1;
    T* hp = (q > 0 && prev >= t0) ? hist + prev * q : nullptr;
    T* fp = (n_frames > 0 && prev >= t0 && prev % field_every == 0)
                ? fld + (prev / field_every) * n : nullptr;
    sampler.before(t);
    if (H.damped)
      hipLaunchKernelGGL((k_step<T, true>), dim3(nblk), dim3(kBlock), 0, st, n, nblk,
                         H.off.as<int64_t>(), H.col.as<int32_t>(), H.val.as<T>(),
                         H.beta.as<T>(), H.gamma.as<T>(), u1, u2, u, dfptr.as<int32_t>(),
                         dfrow.as<int32_t>(), dfalpha.as<T>(), dffrow.as<int32_t>(),
                         dF.as<T>(), frc_rows, static_cast<int32_t>(t), dmptr.as<int32_t>(),
                         dmrow.as<int32_t>(), dmk.as<int32_t>(), hp, fp, field_is_vel, dtT);
    else
      hipLaunchKernelGGL((k_step<T, false>), dim3(nblk), dim3(kBlock), 0, st, n, nblk,
                         H.off.as<int64_t>(), H.col.as<int32_t>(), H.val.as<T>(),
                         static_cast<const T*>(nullptr), static_cast<const T*>(nullptr), u1,
                         u2, u, dfptr.as<int32_t>(), dfrow.as<int32_t>(), dfalpha.as<T>(),
                         dffrow.as<int32_t>(), dF.as<T>(), frc_rows, static_cast<int32_t>(t),
                         dmptr.as<int32_t>(), dmrow.as<int32_t>(), dmk.as<int32_t>(), hp, fp,
                         field_is_vel, dtT);
    sampler.after(t);
    ++launches;
    if (H.nI > 0) {
      hipLaunchKernelGGL((k_iface<T>), dim3(1), dim3(kIfaceThreads), 0, st, H.nI, 1,
                         H.iface.as<int32_t>(), H.cptr.as<int64_t>(), H.ccol.as<int32_t>(),
                         H.cval.as<T>(), static_cast<T>(H.sq), static_cast<T>(H.sr),
                         H.has_xinv ? H.xptr.as<int64_t>() : nullptr, H.xcol.as<int32_t>(),
                         H.xval.as<T>(), u1, u2, u, dy.as<T>(),
                         static_cast<int32_t>(L.fi_k.size()), dfik.as<int32_t>(),
                         dfisrc.as<int32_t>(), dfialpha.as<T>(), dfifrow.as<int32_t>(),
                         dF.as<T>(), frc_rows, static_cast<int32_t>(t));
      ++launches;
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
                         q, dhidx.as<int32_t>(), hp, fp, field_is_vel, dtT);
      ++launches;
    }
  }
  BFE_HIP(hipGetLastError());
  BFE_HIP(hipEventRecord(H.ev1, st));
  BFE_HIP(hipEventSynchronize(H.ev1));
  float ms = 0.f;
  BFE_HIP(hipEventElapsedTime(&ms, H.ev0, H.ev1));
  H.stats.loop_ms = ms;
  H.stats.steps = n_t - t0;
  H.stats.launches = launches;
  sampler.collect();

  std::vector<T> tmp;
  if (q > 0) {
    tmp.resize(static_cast<size_t>(q) * n_t);
    BFE_HIP(hipMemcpy(tmp.data(), hist, dhist.bytes, hipMemcpyDeviceToHost));
    int64_t bad = 0;
    for (size_t i = 0; i < tmp.size(); ++i) {
      hist_out[i] = static_cast<double>(tmp[i]);
      bad += !std::isfinite(hist_out[i]);
    }
    H.stats.nonfinite = bad;
  }
  if (n_frames > 0) {
    tmp.resize(static_cast<size_t>(n) * n_frames);
    BFE_HIP(hipMemcpy(tmp.data(), fld, dfld.bytes, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < tmp.size(); ++i) field_out[i] = static_cast<double>(tmp[i]);
  }
}

template <typename T>
void run_batched(bfe_plan_s& H, int32_t n_t, int32_t n_src, const int32_t* frc_idx,
                 const int32_t* frc_src, int32_t p, const double* frcs, int32_t frc_rows,
                 const int32_t* hist_idx, int32_t q, double* hist_out,
                 int32_t* ti_start_out) {
  const int64_t n = H.n;
  if (n_t < 1) throw std::invalid_argument("n_t must be >= 1");
  if (n_src < 1 || n_src > 64) throw std::invalid_argument("n_src must be in [1, 64]");
  if (p < 0 || q < 0) throw std::invalid_argument("negative count");
  check_indices(frc_idx, p, n, "forcing");
  check_indices(frc_src, p, n_src, "forcing source");
  check_indices(hist_idx, q, n, "history");
  if (p > 0) {
    if (frcs == nullptr) throw std::invalid_argument("forcing functions are NULL");
    if (frc_rows != 1 && frc_rows != p)
      throw std::invalid_argument("forcing functions must have 1 or p rows");
  }
  if (q > 0 && hist_out == nullptr) throw std::invalid_argument("history output is NULL");
  const int32_t S = n_src;
  if (q > 0) std::memset(hist_out, 0, sizeof(double) * static_cast<size_t>(q) * S * n_t);
  H.stats = bfe_run_stats{};
  H.stats.columns = S;
  const int64_t t0 = p > 0 ? first_forced_step(frcs, frc_rows, n_t) : -1;
  if (ti_start_out) *ti_start_out = t0 < 0 ? n_t + 1 : static_cast<int32_t>(t0 + 1);
  if (t0 < 0) return;

  BFE_HIP(hipSetDevice(H.device));
  hipStream_t st = H.stream;
  constexpr int64_t rows_per_block = kBlock / 64;
  const int64_t nblk = (n + rows_per_block - 1) / rows_per_block;
  std::vector<int32_t> fdof(frc_idx, frc_idx + p), fsrc(frc_src, frc_src + p), frow(p);
  for (int32_t i = 0; i < p; ++i) frow[i] = frc_rows == 1 ? 0 : i;
  RunLists L = make_lists(H, rows_per_block, nblk, fdof, fsrc, frow, hist_idx, q);

  std::vector<double> Fh(frcs, frcs + static_cast<size_t>(p > 0 ? frc_rows : 0) * n_t);
  DevBuf dF = upload(cast_vec<T>(Fh), st);
  DevBuf dfptr = upload(L.fptr, st), dfrow = upload(L.f_row, st), dfsrc = upload(L.f_src, st),
         dffrow = upload(L.f_frow, st), dfalpha = upload(cast_vec<T>(L.f_alpha), st);
  DevBuf dfik = upload(L.fi_k, st), dfisrc = upload(L.fi_src, st),
         dfifrow = upload(L.fi_frow, st), dfialpha = upload(cast_vec<T>(L.fi_alpha), st);
  DevBuf dmptr = upload(L.mptr, st), dmrow = upload(L.m_row, st), dmk = upload(L.m_k, st);
  std::vector<int32_t> hidx(hist_idx, hist_idx + q);
  DevBuf dhidx = upload(hidx, st);
  DevBuf dhist(sizeof(T) * static_cast<size_t>(q) * S * n_t);
  DevBuf dy(sizeof(T) * static_cast<size_t>(H.nI) * S);
  DevBuf ubuf(sizeof(T) * static_cast<size_t>(n) * S * 3);
  BFE_HIP(hipMemsetAsync(ubuf.p, 0, ubuf.bytes, st));
  if (dhist.bytes) BFE_HIP(hipMemsetAsync(dhist.p, 0, dhist.bytes, st));
  const int64_t ns = n * S;
  T* U[3] = {ubuf.as<T>(), ubuf.as<T>() + ns, ubuf.as<T>() + 2 * ns};
  T *u = U[0], *u1 = U[1], *u2 = U[2];
  T* hist = dhist.as<T>();
  const int64_t per_t = static_cast<int64_t>(q) * S;
  int64_t launches = 0;
  KernelSampler sampler(H, t0, n_t - t0);

  BFE_HIP(hipEventRecord(H.ev0, st));
  for (int64_t t = t0; t < n_t; ++t) {
    const int64_t prev = t - 1;
    T* hp = (q > 0 && prev >= t0) ? hist + prev * per_t : nullptr;
    sampler.before(t);
    if (H.damped)
      hipLaunchKernelGGL((k_step_multi<T, true>), dim3(nblk), dim3(kBlock), 0, st, n, S, nblk,
                         H.off.as<int64_t>(), H.col.as<int32_t>(), H.val.as<T>(),
                         H.beta.as<T>(), H.gamma.as<T>(), u1, u2, u, dfptr.as<int32_t>(),
                         dfrow.as<int32_t>(), dfsrc.as<int32_t>(), dfalpha.as<T>(),
                         dffrow.as<int32_t>(), dF.as<T>(), frc_rows, static_cast<int32_t>(t),
                         dmptr.as<int32_t>(), dmrow.as<int32_t>(), dmk.as<int32_t>(), q, hp);
    else
      hipLaunchKernelGGL((k_step_multi<T, false>), dim3(nblk), dim3(kBlock), 0, st, n, S, nblk,
                         H.off.as<int64_t>(), H.col.as<int32_t>(), H.val.as<T>(),
                         static_cast<const T*>(nullptr), static_cast<const T*>(nullptr), u1,
                         u2, u, dfptr.as<int32_t>(), dfrow.as<int32_t>(), dfsrc.as<int32_t>(),
                         dfalpha.as<T>(), dffrow.as<int32_t>(), dF.as<T>(), frc_rows,
                         static_cast<int32_t>(t), dmptr.as<int32_t>(), dmrow.as<int32_t>(),
                         dmk.as<int32_t>(), q, hp);
    sampler.after(t);
    ++launches;
    if (H.nI > 0) {
      hipLaunchKernelGGL((k_iface<T>), dim3(1), dim3(kIfaceThreads), 0, st, H.nI, S,
                         H.iface.as<int32_t>(), H.cptr.as<int64_t>(), H.ccol.as<int32_t>(),
                         H.cval.as<T>(), static_cast<T>(H.sq), static_cast<T>(H.sr),
                         H.has_xinv ? H.xptr.as<int64_t>() : nullptr, H.xcol.as<int32_t>(),
                         H.xval.as<T>(), u1, u2, u, dy.as<T>(),
                         static_cast<int32_t>(L.fi_k.size()), dfik.as<int32_t>(),
                         dfisrc.as<int32_t>(), dfialpha.as<T>(), dfifrow.as<int32_t>(),
                         dF.as<T>(), frc_rows, static_cast<int32_t>(t));
      ++launches;
    }
    T* tmp = u2; u2 = u1; u1 = u; u = tmp;
  }
  if (q > 0) {
    const int64_t work = per_t;
    hipLaunchKernelGGL((k_snap<T>), dim3((work + 255) / 256), dim3(256), 0, st, n, S, u1, u2, q,
                       dhidx.as<int32_t>(), hist + (n_t - 1) * per_t, static_cast<T*>(nullptr),
                       0, static_cast<T>(H.dt));
    ++launches;
  }
  BFE_HIP(hipGetLastError());
  BFE_HIP(hipEventRecord(H.ev1, st));
  BFE_HIP(hipEventSynchronize(H.ev1));
  float ms = 0.f;
  BFE_HIP(hipEventElapsedTime(&ms, H.ev0, H.ev1));
  H.stats.loop_ms = ms;
  H.stats.steps = n_t - t0;
  H.stats.launches = launches;
  sampler.collect();
  if (q > 0) {
    std::vector<T> tmp(static_cast<size_t>(per_t) * n_t);
    BFE_HIP(hipMemcpy(tmp.data(), hist, dhist.bytes, hipMemcpyDeviceToHost));
    int64_t bad = 0;
    for (size_t i = 0; i < tmp.size(); ++i) {
      hist_out[i] = static_cast<double>(tmp[i]);
      bad += !std::isfinite(hist_out[i]);
    }
    H.stats.nonfinite = bad;
  }
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
    bfe::Csr K = bfe::normalise_csr(n, k_rowptr, k_col, k_val, "K");
    bfe::Csr C = bfe::normalise_csr(n, c_rowptr, c_col, c_val, "C");
    bfe::Plan P = bfe::build_plan(n, K, m, C, dt, mode);

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
    H->n_slices = P.op.n_slices;
    H->nnz_sell = P.op.off.back();
    H->nnz_k = P.nnz_k;
    H->nI = static_cast<int32_t>(P.iface.size());
    H->nnz_coff = P.coff.nnz();
    H->nnz_xinv = P.has_xinv ? P.xinv.nnz() : 0;
    H->sq = P.sq;
    H->sr = P.sr;
    H->has_xinv = P.has_xinv;
    H->alpha = P.alpha;
    H->iface_pos = P.iface_pos;
    if (precision == BFE_PREC_DOUBLE) upload_plan<double>(*H, P);
    else upload_plan<float>(*H, P);
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
                         field_every, field_is_vel, field_out, ti_start_out);
    else
      run_single<float>(*h, n_t, frc_idx, p, frcs, frc_rows, hist_idx, q, hist_out,
                        field_every, field_is_vel, field_out, ti_start_out);
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
    I.device_bytes = static_cast<int64_t>(
        h->off.bytes + h->col.bytes + h->val.bytes + h->beta.bytes + h->gamma.bytes +
        h->iface.bytes + h->cptr.bytes + h->ccol.bytes + h->cval.bytes + h->xptr.bytes +
        h->xcol.bytes + h->xval.bytes);
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

// Subdomain coupling operations on the GPU (FP64):
//
//   bfe_convolve         code/internal/fn_convolve.m:22-34: each signal f
//                        (n_t samples) convolved with g through FFTs of
//                        length 2 n_t, the first n_t samples kept. hipFFT
//                        (D2Z, Z2D) in batches, the spectra multiplied (and
//                        scaled by 1/(2 n_t)) by a kernel of this file.
//   bfe_disps_to_forces  code/fn_convert_disps_to_forces_v2.m:1-29: the
//                        forces M a + C v + K d on the rows of layers 2
//                        and 3 that reproduce a displacement history, with
//                        the frozen layers' displacements set to zero; one
//                        thread per (row, time sample).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "bfe.h"

namespace bfe {
void set_subdomain_error(const std::string& msg);
}

namespace {

struct sub_failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define SUB_HIP(expr)                                                           \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      throw sub_failure(std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define SUB_FFT(expr)                                                           \
  do {                                                                          \
    hipfftResult r_ = (expr);                                                   \
    if (r_ != HIPFFT_SUCCESS)                                                   \
      throw sub_failure(std::string(#expr) + ": hipFFT status " + std::to_string(r_)); \
  } while (0)

struct Mem {
  void* p = nullptr;
  explicit Mem(size_t b) {
    if (b && hipMalloc(&p, b) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw std::bad_alloc();
    }
  }
  Mem(const Mem&) = delete;
  Mem& operator=(const Mem&) = delete;
  ~Mem() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

struct FftPlan {
  hipfftHandle h = 0;
  bool ok = false;
  FftPlan(int P, int batch, hipfftType type) {
    int n[1] = {P};
    const int half = P / 2 + 1;
    const bool fwd = type == HIPFFT_D2Z;
    SUB_FFT(hipfftPlanMany(&h, 1, n, nullptr, 1, fwd ? P : half, nullptr, 1, fwd ? half : P, type,
                           batch));
    ok = true;
  }
  FftPlan(const FftPlan&) = delete;
  FftPlan& operator=(const FftPlan&) = delete;
  ~FftPlan() {
    if (ok) (void)hipfftDestroy(h);
  }
};

// Zero-padded copy: dst (rows x P) <- src (rows x n, row-major), columns
// n..P-1 zero (n may exceed P: truncation, as fft(g, P) does).
__global__ void k_pad(int64_t rows, int32_t n, int32_t P, const double* __restrict__ src,
                      double* __restrict__ dst) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows * P) return;
  const int64_t r = i / P;
  const int32_t k = static_cast<int32_t>(i - r * P);
  dst[i] = k < n ? src[r * n + k] : 0.0;
}

// F[s][k] *= G[s or 0][k] / P (complex), k < half.
__global__ void k_cmul(int64_t rows, int32_t half, double2* __restrict__ F,
                       const double2* __restrict__ G, int32_t g_per_signal, double inv_p) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows * half) return;
  const int64_t s = i / half;
  const int64_t k = i - s * half;
  const double2 a = F[i];
  const double2 b = G[(g_per_signal ? s * half : 0) + k];
  F[i] = make_double2((a.x * b.x - a.y * b.y) * inv_p, (a.x * b.y + a.y * b.x) * inv_p);
}

// dst (rows x n) <- first n columns of src (rows x P).
__global__ void k_crop(int64_t rows, int32_t n, int32_t P, const double* __restrict__ src,
                       double* __restrict__ dst) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows * n) return;
  const int64_t r = i / n;
  dst[i] = src[r * P + (i - r * n)];
}

inline unsigned grid(int64_t n, int t = 256) { return static_cast<unsigned>((n + t - 1) / t); }

// Forces of fn_convert_disps_to_forces_v2 on force rows fr[0..nf) at every
// time sample: D is the displacement history (n x n_t, column-major) with
// the frozen rows read as zero (layers 3, 4 for in_or_out = 0 'in', layers
// 1, 2 for 1 'out', :5-10), d_t for t < 0 zero. For output column t:
//   a = (d_t - 2 d_{t-1} + d_{t-2}) / dt^2
//   v = (d_t - d_{t-2}) / (2 dt)     (velocity at the current step)
//     = (d_{t-1} - d_{t-2}) / dt     (at the last half step)
//   f = m_r a_r + sum_j C_rj v_j + sum_j K_rj d_{t-1, j}
#pragma clang fp contract(off)
__global__ void k_forces(int64_t n, int32_t n_t, int64_t nf, const int32_t* __restrict__ fr,
                         const int64_t* __restrict__ kp, const int32_t* __restrict__ kc,
                         const double* __restrict__ kv, const int64_t* __restrict__ cp,
                         const int32_t* __restrict__ cc, const double* __restrict__ cv,
                         const double* __restrict__ m, const int32_t* __restrict__ lyrs,
                         int32_t in_or_out, const double* __restrict__ D, double dt,
                         int32_t mode, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nf * n_t) return;
  const int64_t fi = i % nf;
  const int32_t t = static_cast<int32_t>(i / nf);
  const int64_t r = fr[fi];
  const int32_t lo = in_or_out == 0 ? 3 : 1;  // frozen layers lo and lo + 1
  auto d = [&](int64_t j, int32_t tt) -> double {
    const int32_t l = lyrs[j];
    return (tt < 0 || l == lo || l == lo + 1) ? 0.0 : D[static_cast<int64_t>(tt) * n + j];
  };
  const double dt2 = dt * dt;
  const double a = (d(r, t) - 2.0 * d(r, t - 1) + d(r, t - 2)) / dt2;
  double sc = 0.0;
  for (int64_t q = cp[r]; q < cp[r + 1]; ++q) {
    const int64_t j = cc[q];
    const double v = mode == BFE_MODE_VEL_AT_CURRENT ? (d(j, t) - d(j, t - 2)) / (2.0 * dt)
                               : (d(j, t - 1) - d(j, t - 2)) / dt;
    sc += cv[q] * v;
  }
  double sk = 0.0;
  for (int64_t q = kp[r]; q < kp[r + 1]; ++q) sk += kv[q] * d(kc[q], t - 1);
  out[static_cast<int64_t>(t) * nf + fi] = m[r] * a + sc + sk;
}
#pragma clang fp contract(on)

// Device memory of bfe_disps_to_forces_noalloc: every device array of that
// entry point (its staged inputs and its output) is carved from this one
// static array, so the call allocates nothing. Each GPU holds its own copy
// from the library's load until the process ends. In host code `scratch`
// only names the symbol: its device address comes from hipGetSymbolAddress,
// once per device (scratch_base).
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_base[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current).
void* scratch_base(int32_t device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_base[device] == nullptr) {
    void* p = nullptr;
    SUB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) throw sub_failure("hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_base[device] = p;
  }
  return g_scratch_base[device];
}

// 256-byte aligned sub-arrays of one scratch region, handed out in order.
struct ScratchCarve {
  size_t used = 0;
  unsigned char* base = nullptr;
  size_t take(size_t bytes) {
    const size_t off = used;
    used += (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
    return off;
  }
  template <typename T>
  T* at(size_t off) const { return reinterpret_cast<T*>(base + off); }
};

}  // namespace

extern "C" {

bfe_status bfe_convolve(int32_t device, int64_t n_sig, int32_t n_t, const double* f, int32_t n_g,
                        const double* g, int32_t g_per_signal, double* h) {
  try {
    if (n_sig < 0 || n_t < 0 || n_g < 0) throw std::invalid_argument("negative size");
    if (n_sig == 0 || n_t == 0) return BFE_OK;
    if (!f || !h || (n_g > 0 && !g)) throw std::invalid_argument("NULL array");
    if (n_t > (1 << 28)) throw std::invalid_argument("signal too long");
    SUB_HIP(hipSetDevice(device));
    const int32_t P = 2 * n_t;  // fn_convolve.m: fft_pts = 2 * size(f, dim)
    const int32_t half = P / 2 + 1;
    // signals per batch: about 512 MB of work arrays
    const int64_t per = static_cast<int64_t>(P) * 8 + static_cast<int64_t>(half) * 16 * 2;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n_sig, (512ll << 20) / per));
    const int32_t ng = std::min(n_g, P);
    Mem G(sizeof(double2) * half * (g_per_signal ? chunk : 1));
    Mem rin(sizeof(double) * P * chunk), spec(sizeof(double2) * half * chunk);
    Mem src(sizeof(double) * std::max<int64_t>(n_t, n_g) * chunk);
    Mem out(sizeof(double) * n_t * chunk);
    if (!g_per_signal) {  // one spectrum of g for every signal
      FftPlan pg(P, 1, HIPFFT_D2Z);
      if (n_g > 0) SUB_HIP(hipMemcpy(src.p, g, sizeof(double) * n_g, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_pad, dim3(grid(P)), dim3(256), 0, nullptr, int64_t(1), ng, P,
                         src.as<double>(), rin.as<double>());
      SUB_HIP(hipGetLastError());
      SUB_FFT(hipfftExecD2Z(pg.h, rin.as<hipfftDoubleReal>(), G.as<hipfftDoubleComplex>()));
    }
    for (int64_t s0 = 0; s0 < n_sig; s0 += chunk) {
      const int64_t b = std::min(chunk, n_sig - s0);
      FftPlan pf(P, static_cast<int>(b), HIPFFT_D2Z);
      FftPlan pi(P, static_cast<int>(b), HIPFFT_Z2D);
      if (g_per_signal) {
        if (n_g > 0)
          SUB_HIP(hipMemcpy(src.p, g + s0 * n_g, sizeof(double) * n_g * b, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_pad, dim3(grid(b * P)), dim3(256), 0, nullptr, b, n_g, P,
                           src.as<double>(), rin.as<double>());
        SUB_HIP(hipGetLastError());
        SUB_FFT(hipfftExecD2Z(pf.h, rin.as<hipfftDoubleReal>(), G.as<hipfftDoubleComplex>()));
      }
      SUB_HIP(hipMemcpy(src.p, f + s0 * n_t, sizeof(double) * n_t * b, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_pad, dim3(grid(b * P)), dim3(256), 0, nullptr, b, n_t, P,
                         src.as<double>(), rin.as<double>());
      SUB_HIP(hipGetLastError());
      SUB_FFT(hipfftExecD2Z(pf.h, rin.as<hipfftDoubleReal>(), spec.as<hipfftDoubleComplex>()));
      hipLaunchKernelGGL(k_cmul, dim3(grid(b * half)), dim3(256), 0, nullptr, b, half,
                         spec.as<double2>(), G.as<double2>(), g_per_signal,
                         1.0 / static_cast<double>(P));
      SUB_HIP(hipGetLastError());
      SUB_FFT(hipfftExecZ2D(pi.h, spec.as<hipfftDoubleComplex>(), rin.as<hipfftDoubleReal>()));
      hipLaunchKernelGGL(k_crop, dim3(grid(b * n_t)), dim3(256), 0, nullptr, b, n_t, P,
                         rin.as<double>(), out.as<double>());
      SUB_HIP(hipGetLastError());
      SUB_HIP(hipMemcpy(h + s0 * n_t, out.p, sizeof(double) * n_t * b, hipMemcpyDeviceToHost));
    }
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    bfe::set_subdomain_error(std::string("invalid argument: ") + e.what());
    return BFE_ERR_ARG;
  } catch (const sub_failure& e) {
    (void)hipGetLastError();
    bfe::set_subdomain_error(std::string("HIP error: ") + e.what());
    return BFE_ERR_HIP;
  } catch (const std::bad_alloc&) {
    bfe::set_subdomain_error("out of memory");
    return BFE_ERR_NOMEM;
  } catch (const std::exception& e) {
    bfe::set_subdomain_error(std::string("internal error: ") + e.what());
    return BFE_ERR_INTERNAL;
  }
}

bfe_status bfe_disps_to_forces(int32_t device, int64_t n, const int64_t* k_rowptr,
                               const int32_t* k_col, const double* k_val,
                               const int64_t* c_rowptr, const int32_t* c_col,
                               const double* c_val, const double* m, double dt, int32_t n_t,
                               const double* disp, const int32_t* lyrs, int32_t in_or_out,
                               int32_t solver_mode, int64_t n_force, int32_t* force_idx,
                               double* force) {
  try {
    if (n <= 0 || n > INT32_MAX || n_t < 0) throw std::invalid_argument("bad sizes");
    if (!k_rowptr || !k_col || !k_val || !c_rowptr || !c_col || !c_val || !m || !lyrs)
      throw std::invalid_argument("NULL array");
    if (in_or_out != 0 && in_or_out != 1) throw std::invalid_argument("in_or_out must be 0 or 1");
    if (solver_mode != 0 && solver_mode != 1) throw std::invalid_argument("bad solver mode");
    if (!(dt > 0.0)) throw std::invalid_argument("time step must be positive");
    // force set: layers 2 and 3, ascending (fn_convert_disps_to_forces_v2.m:2-4);
    // frozen: layers 3, 4 ('in') or 1, 2 ('out') (:5-10)
    std::vector<int32_t> fr;
    for (int64_t j = 0; j < n; ++j) {
      const int32_t l = lyrs[j];
      if (l == 2 || l == 3) fr.push_back(static_cast<int32_t>(j));
    }
    if (static_cast<int64_t>(fr.size()) != n_force)
      throw std::invalid_argument("n_force must be the number of DOFs in layers 2 and 3 (" +
                                  std::to_string(fr.size()) + ")");
    if (n_force == 0) return BFE_OK;
    if (!force_idx || !force) throw std::invalid_argument("NULL output");
    std::memcpy(force_idx, fr.data(), sizeof(int32_t) * fr.size());
    if (n_t == 0) return BFE_OK;
    if (!disp) throw std::invalid_argument("NULL displacements");
    const int64_t nk = k_rowptr[n], nc = c_rowptr[n];
    for (int64_t q = 0; q < nk; ++q)
      if (k_col[q] < 0 || k_col[q] >= n) throw std::invalid_argument("K column out of range");
    for (int64_t q = 0; q < nc; ++q)
      if (c_col[q] < 0 || c_col[q] >= n) throw std::invalid_argument("C column out of range");
    SUB_HIP(hipSetDevice(device));
    auto up = [](const void* h, size_t bytes) {
      auto b = std::make_unique<Mem>(bytes);
      if (bytes) SUB_HIP(hipMemcpy(b->p, h, bytes, hipMemcpyHostToDevice));
      return b;
    };
    auto dkp = up(k_rowptr, sizeof(int64_t) * (n + 1));
    auto dkc = up(k_col, sizeof(int32_t) * std::max<int64_t>(nk, 1));
    auto dkv = up(k_val, sizeof(double) * std::max<int64_t>(nk, 1));
    auto dcp = up(c_rowptr, sizeof(int64_t) * (n + 1));
    auto dcc = up(c_col, sizeof(int32_t) * std::max<int64_t>(nc, 1));
    auto dcv = up(c_val, sizeof(double) * std::max<int64_t>(nc, 1));
    auto dm = up(m, sizeof(double) * n);
    auto dly = up(lyrs, sizeof(int32_t) * n);
    auto dfr = up(fr.data(), sizeof(int32_t) * fr.size());
    auto dD = up(disp, sizeof(double) * static_cast<size_t>(n) * n_t);
    Mem dout(sizeof(double) * static_cast<size_t>(n_force) * n_t);
    hipLaunchKernelGGL(k_forces, dim3(grid(n_force * n_t)), dim3(256), 0, nullptr, n, n_t,
                       n_force, dfr->as<int32_t>(), dkp->as<int64_t>(), dkc->as<int32_t>(),
                       dkv->as<double>(), dcp->as<int64_t>(), dcc->as<int32_t>(),
                       dcv->as<double>(), dm->as<double>(), dly->as<int32_t>(), in_or_out,
                       dD->as<double>(), dt, solver_mode, dout.as<double>());
    SUB_HIP(hipGetLastError());
    SUB_HIP(hipMemcpy(force, dout.p, sizeof(double) * static_cast<size_t>(n_force) * n_t,
                      hipMemcpyDeviceToHost));
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    bfe::set_subdomain_error(std::string("invalid argument: ") + e.what());
    return BFE_ERR_ARG;
  } catch (const sub_failure& e) {
    (void)hipGetLastError();
    bfe::set_subdomain_error(std::string("HIP error: ") + e.what());
    return BFE_ERR_HIP;
  } catch (const std::bad_alloc&) {
    bfe::set_subdomain_error("out of memory");
    return BFE_ERR_NOMEM;
  } catch (const std::exception& e) {
    bfe::set_subdomain_error(std::string("internal error: ") + e.what());
    return BFE_ERR_INTERNAL;
  }
}

bfe_status bfe_disps_to_forces_noalloc(int32_t device, int64_t n, const int64_t* k_rowptr,
                                       const int32_t* k_col, const double* k_val,
                                       const int64_t* c_rowptr, const int32_t* c_col,
                                       const double* c_val, const double* m, double dt,
                                       int32_t n_t, const double* disp, const int32_t* lyrs,
                                       int32_t in_or_out, int32_t solver_mode, int64_t n_force,
                                       int32_t* force_idx, double* force) {
  try {
    if (n <= 0 || n > INT32_MAX || n_t < 0) throw std::invalid_argument("bad sizes");
    if (!k_rowptr || !k_col || !k_val || !c_rowptr || !c_col || !c_val || !m || !lyrs)
      throw std::invalid_argument("NULL array");
    if (in_or_out != 0 && in_or_out != 1) throw std::invalid_argument("in_or_out must be 0 or 1");
    if (solver_mode != 0 && solver_mode != 1) throw std::invalid_argument("bad solver mode");
    if (!(dt > 0.0)) throw std::invalid_argument("time step must be positive");
    // force set: layers 2 and 3, ascending (fn_convert_disps_to_forces_v2.m:2-4),
    // counted first so that force_idx is written only when it fits
    int64_t nf = 0;
    for (int64_t j = 0; j < n; ++j) nf += (lyrs[j] == 2 || lyrs[j] == 3);
    if (nf != n_force)
      throw std::invalid_argument("n_force must be the number of DOFs in layers 2 and 3");
    if (n_force == 0) return BFE_OK;
    if (!force_idx || !force) throw std::invalid_argument("NULL output");
    for (int64_t j = 0, k = 0; j < n; ++j)
      if (lyrs[j] == 2 || lyrs[j] == 3) force_idx[k++] = static_cast<int32_t>(j);
    if (n_t == 0) return BFE_OK;
    if (!disp) throw std::invalid_argument("NULL displacements");
    const int64_t nk = k_rowptr[n], nc = c_rowptr[n];
    if (nk < 0 || nc < 0) throw std::invalid_argument("bad row pointers");
    for (int64_t q = 0; q < nk; ++q)
      if (k_col[q] < 0 || k_col[q] >= n) throw std::invalid_argument("K column out of range");
    for (int64_t q = 0; q < nc; ++q)
      if (c_col[q] < 0 || c_col[q] >= n) throw std::invalid_argument("C column out of range");
    // the layout in the scratch array, checked against its size before any
    // device call (first in floating point, so the byte counts cannot wrap)
    const double need = 8.0 * (4.0 * n + 2.0 * (nk + nc) + static_cast<double>(n) * n_t +
                               static_cast<double>(n_force) * n_t);
    if (need > static_cast<double>(kScratchBytes))
      throw std::invalid_argument("inputs and forces need more than the " +
                                  std::to_string(kScratchBytes) + " bytes of device memory that "
                                  "bfe_disps_to_forces_noalloc holds (use bfe_disps_to_forces)");
    ScratchCarve sc;
    const size_t o_kp = sc.take(sizeof(int64_t) * (n + 1));
    const size_t o_kc = sc.take(sizeof(int32_t) * nk);
    const size_t o_kv = sc.take(sizeof(double) * nk);
    const size_t o_cp = sc.take(sizeof(int64_t) * (n + 1));
    const size_t o_cc = sc.take(sizeof(int32_t) * nc);
    const size_t o_cv = sc.take(sizeof(double) * nc);
    const size_t o_m = sc.take(sizeof(double) * n);
    const size_t o_ly = sc.take(sizeof(int32_t) * n);
    const size_t o_fr = sc.take(sizeof(int32_t) * n_force);
    const size_t o_d = sc.take(sizeof(double) * static_cast<size_t>(n) * n_t);
    const size_t o_out = sc.take(sizeof(double) * static_cast<size_t>(n_force) * n_t);
    if (sc.used > kScratchBytes)
      throw std::invalid_argument("inputs and forces need " + std::to_string(sc.used) +
                                  " bytes of device memory; bfe_disps_to_forces_noalloc holds at "
                                  "most " + std::to_string(kScratchBytes) +
                                  " (use bfe_disps_to_forces)");
    SUB_HIP(hipSetDevice(device));
    sc.base = static_cast<unsigned char*>(scratch_base(device));
    auto put = [&](size_t off, const void* h, size_t bytes) {
      if (bytes) SUB_HIP(hipMemcpy(sc.base + off, h, bytes, hipMemcpyHostToDevice));
    };
    put(o_kp, k_rowptr, sizeof(int64_t) * (n + 1));
    put(o_kc, k_col, sizeof(int32_t) * nk);
    put(o_kv, k_val, sizeof(double) * nk);
    put(o_cp, c_rowptr, sizeof(int64_t) * (n + 1));
    put(o_cc, c_col, sizeof(int32_t) * nc);
    put(o_cv, c_val, sizeof(double) * nc);
    put(o_m, m, sizeof(double) * n);
    put(o_ly, lyrs, sizeof(int32_t) * n);
    put(o_fr, force_idx, sizeof(int32_t) * n_force);
    put(o_d, disp, sizeof(double) * static_cast<size_t>(n) * n_t);
    // every output element is written by the kernel: nothing to clear
    hipLaunchKernelGGL(k_forces, dim3(grid(n_force * n_t)), dim3(256), 0, nullptr, n, n_t,
                       n_force, sc.at<int32_t>(o_fr), sc.at<int64_t>(o_kp), sc.at<int32_t>(o_kc),
                       sc.at<double>(o_kv), sc.at<int64_t>(o_cp), sc.at<int32_t>(o_cc),
                       sc.at<double>(o_cv), sc.at<double>(o_m), sc.at<int32_t>(o_ly), in_or_out,
                       sc.at<double>(o_d), dt, solver_mode, sc.at<double>(o_out));
    SUB_HIP(hipGetLastError());
    SUB_HIP(hipMemcpy(force, sc.base + o_out, sizeof(double) * static_cast<size_t>(n_force) * n_t,
                      hipMemcpyDeviceToHost));
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    bfe::set_subdomain_error(std::string("invalid argument: ") + e.what());
    return BFE_ERR_ARG;
  } catch (const sub_failure& e) {
    (void)hipGetLastError();
    bfe::set_subdomain_error(std::string("HIP error: ") + e.what());
    return BFE_ERR_HIP;
  } catch (const std::exception& e) {
    bfe::set_subdomain_error(std::string("internal error: ") + e.what());
    return BFE_ERR_INTERNAL;
  }
}

}  // extern "C"

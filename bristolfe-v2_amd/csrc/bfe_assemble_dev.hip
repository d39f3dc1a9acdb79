// Device-side global assembly (GPU, FP64).
//
// Same result as code/internal/fn_build_global_matrices_v5.m:50-159 with
// fn_reduce_global_matrices.m and fn_create_fast_lookup.m, for meshes of
// CPE3, AC2D3, C3D4 and ASI2D2 (fluid-solid interface) elements: element
// matrices with the absorbing modifications (:115-116), global index
// (node-1)*max_df + dof (:127-128), duplicates summed (:145-147), DOFs whose
// row/column is empty in K, M and C removed in order, and gl_lookup.
//
// Everything after the upload runs on the GPU:
//   1. node -> element incidence: counts (atomics), an exclusive scan, a fill
//      and a per-node sort, so every node lists its elements in ascending
//      order whatever order the atomics took;
//   2. pass A, one thread per node: the node's rows of K and C are gathered
//      from the rows of its elements' matrices (computed in registers, never
//      stored), summed per column in ascending element order, and counted;
//   3. scans of the per-row flags and counts give the compacted DOF numbers
//      and both CSR row pointers;
//   4. pass B repeats the gather and writes the columns (ascending) and
//      values.
// The element formulas are those of the host assembler (bfe_assemble.cpp)
// in the same order of operations, with floating-point contraction off, so
// both give the same values up to the last bit of exp/pow in absorbing
// elements. M is diagonal for these element types and comes back as a
// vector; C carries the absorbing diagonal and the ASI2D2 couplings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bfe.h"

namespace bfe {
void set_assembly_error(const std::string& msg);  // bfe_assemble.cpp
}

namespace {

enum ElType : int32_t { kCPE3 = 0, kAC2D3 = 1, kC3D4 = 2, kASI2D2 = 3 };
constexpr int kNbMax = 40;  // distinct nodes around one node, itself included
constexpr int kMaxDf = 4;   // DOFs per node: u1, u2, u3, p
constexpr int kThreads = 128;
constexpr int kScanBlock = 1024;  // elements per workgroup of the scan

struct dev_failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define ASM_HIP(expr)                                                           \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      throw dev_failure(std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

// Device allocation owned by the host object.
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  Buf() = default;
  explicit Buf(size_t b) : bytes(b) {
    if (b == 0) return;
    if (hipMalloc(&p, b) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw std::bad_alloc();
    }
  }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  Buf& operator=(Buf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~Buf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

template <typename T>
Buf upload(const T* h, size_t count) {
  Buf b(sizeof(T) * count);
  if (count) ASM_HIP(hipMemcpy(b.p, h, sizeof(T) * count, hipMemcpyHostToDevice));
  return b;
}

// Mesh as the kernels see it (device pointers, scalars).
struct DMesh {
  int64_t n_nodes = 0, n_el = 0;
  int32_t dim = 0, nper = 0, mask = 0, max_df = 0;
  double p_law = 3.0, max_damp = 3.1415e7, log_msr = 0.0;
  const double* nds = nullptr;
  const int32_t* els = nullptr;
  const int32_t* typ = nullptr;
  const int32_t* mat = nullptr;
  const double* abs_i = nullptr;
  const double* D = nullptr;
  const double* rho = nullptr;
  const int64_t* inc_ptr = nullptr;
  const int32_t* inc = nullptr;
};

__host__ __device__ inline int nodes_of(int32_t t) {
  return t == kC3D4 ? 4 : (t == kASI2D2 ? 2 : 3);
}
// DOF d (1-based) of type t before the dof_to_use filter
__device__ inline bool type_has_dof(int32_t t, int d) {
  if (t == kAC2D3) return d == 4;
  if (t == kASI2D2) return d >= 1 && d <= 4;
  return d >= 1 && d <= 3;
}
__host__ __device__ inline bool keep_dof(int32_t mask, int d) {
  return mask == 0 || ((mask >> (d - 1)) & 1);
}

// ---------------------------------------------------------------------------
// Element rows. Row (a, d) of element e: add(b, d2, k, c) for every entry of
// that row of K_e (times the stiffness reduction) and of C_e whose column
// DOF d2 is kept; *md and *cd receive the lumped mass of (a, d) and its
// absorbing damping. Closed forms: CPE3 K_e = (J/2) B'DB with engineering
// strains (e11, e22, e33, g23, g13, g12) and d/dz = 0, M = J rho / 6; AC2D3
// K_e = -(b b' + c c') / (2 J rho), M = -J / (6 K); C3D4 K_e = (J/6) B'DB,
// M = J rho / 24; ASI2D2 couples u1 (y2/4 - y1/4) and u2 (x1/4 - x2/4) of
// either node with p of either node, in C only.
#pragma clang fp contract(off)
template <typename Add>
__device__ void element_row(const DMesh& m, int64_t e, int a, int d, Add&& add, double* md,
                            double* cd) {
  const int32_t t = m.typ[e];
  const int32_t* v = m.els + e * m.nper;
  *md = 0.0;
  *cd = 0.0;
  auto X = [&](int k, int c) { return m.nds[static_cast<int64_t>(v[k]) * m.dim + c]; };
  if (t == kASI2D2) {
    const double x1 = X(0, 0), y1 = X(0, 1), x2 = X(1, 0), y2 = X(1, 1);
    const double cx = y2 / 4.0 - y1 / 4.0;
    const double cy = x1 / 4.0 - x2 / 4.0;
    if (d == 1 || d == 2) {
      if (keep_dof(m.mask, 4))
        for (int b = 0; b < 2; ++b) add(b, 4, 0.0, d == 1 ? cx : cy);
    } else if (d == 4) {
      for (int b = 0; b < 2; ++b) {
        if (keep_dof(m.mask, 1)) add(b, 1, 0.0, cx);
        if (keep_dof(m.mask, 2)) add(b, 2, 0.0, cy);
      }
    }
    return;
  }
  const double* D = m.D + static_cast<int64_t>(m.mat[e]) * 36;
  const double rho = m.rho[m.mat[e]];
  const double ab = m.abs_i != nullptr ? m.abs_i[e] : 0.0;
  const double damp = ab != 0.0 ? pow(ab, m.p_law) * m.max_damp : 0.0;
  const double red = ab != 0.0 ? exp(m.log_msr * pow(ab, m.p_law + 1.0)) : 1.0;
  if (t == kCPE3 || t == kAC2D3) {
    const double x1 = X(0, 0), y1 = X(0, 1), x2 = X(1, 0), y2 = X(1, 1), x3 = X(2, 0),
                 y3 = X(2, 1);
    const double J = x1 * y2 - y1 * x2 - x1 * y3 + y1 * x3 + x2 * y3 - y2 * x3;
    const double b[3] = {y2 - y3, y3 - y1, y1 - y2};
    const double c[3] = {x3 - x2, x1 - x3, x2 - x1};
    if (t == kAC2D3) {
      for (int j = 0; j < 3; ++j) {
        const double kv = -(b[a] * b[j] + c[a] * c[j]) / (2.0 * J * rho);
        if (kv != 0.0) add(j, 4, kv * red, 0.0);
      }
      *md = -J / (6.0 * D[0]);
      *cd = *md * damp;
      return;
    }
    double B[6][9];
    for (int r = 0; r < 6; ++r)
      for (int j = 0; j < 9; ++j) B[r][j] = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double bi = b[i] / J, ci = c[i] / J;
      B[0][3 * i] = bi;
      B[1][3 * i + 1] = ci;
      B[3][3 * i + 2] = ci;
      B[4][3 * i + 2] = bi;
      B[5][3 * i] = ci;
      B[5][3 * i + 1] = bi;
    }
    const int i = 3 * a + (d - 1);
    for (int j = 0; j < 9; ++j) {
      const int d2 = j % 3 + 1;
      if (!keep_dof(m.mask, d2)) continue;
      double s = 0.0;
      for (int r = 0; r < 6; ++r) {
        double db = 0.0;
        for (int q = 0; q < 6; ++q) db += D[r * 6 + q] * B[q][j];
        s += B[r][i] * db;
      }
      const double kv = s * (J / 2.0);
      if (kv != 0.0) add(j / 3, d2, kv * red, 0.0);
    }
    *md = J * rho / 6.0;
    *cd = *md * damp;
    return;
  }
  // C3D4
  double E[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) E[r][c] = X(r + 1, c) - X(0, c);
  const double J = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) -
                   E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                   E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
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
  for (int p = 0; p < 4; ++p)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += gref[p][k] * inv[j][k];
      G[p][j] = s;
    }
  double B[6][12];
  for (int r = 0; r < 6; ++r)
    for (int j = 0; j < 12; ++j) B[r][j] = 0.0;
  for (int p = 0; p < 4; ++p) {
    B[0][3 * p] = G[p][0];
    B[1][3 * p + 1] = G[p][1];
    B[2][3 * p + 2] = G[p][2];
    B[3][3 * p + 1] = G[p][2];
    B[3][3 * p + 2] = G[p][1];
    B[4][3 * p] = G[p][2];
    B[4][3 * p + 2] = G[p][0];
    B[5][3 * p] = G[p][1];
    B[5][3 * p + 1] = G[p][0];
  }
  const int i = 3 * a + (d - 1);
  for (int j = 0; j < 12; ++j) {
    const int d2 = j % 3 + 1;
    if (!keep_dof(m.mask, d2)) continue;
    double s = 0.0;
    for (int r = 0; r < 6; ++r) {
      double db = 0.0;
      for (int q = 0; q < 6; ++q) db += D[r * 6 + q] * B[q][j];
      s += B[r][i] * db;
    }
    const double kv = s * (J / 6.0);
    if (kv != 0.0) add(j / 3, d2, kv * red, 0.0);
  }
  *md = J * rho / 24.0;
  *cd = *md * damp;
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Incidence

__global__ void k_inc_count(DMesh m, int64_t* cnt) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m.n_el) return;
  const int nn = nodes_of(m.typ[e]);
  for (int k = 0; k < nn; ++k)
    atomicAdd(reinterpret_cast<unsigned long long*>(cnt + m.els[e * m.nper + k]), 1ull);
}

__global__ void k_inc_fill(DMesh m, const int64_t* ptr, int32_t* cursor, int32_t* inc) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m.n_el) return;
  const int nn = nodes_of(m.typ[e]);
  for (int k = 0; k < nn; ++k) {
    const int32_t v = m.els[e * m.nper + k];
    const int32_t pos = atomicAdd(cursor + v, 1);
    inc[ptr[v] + pos] = static_cast<int32_t>(e);
  }
}

// Each node's elements in ascending order (insertion sort of its segment).
__global__ void k_inc_sort(int64_t n_nodes, const int64_t* ptr, int32_t* inc) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n_nodes) return;
  const int64_t b = ptr[v], e = ptr[v + 1];
  for (int64_t i = b + 1; i < e; ++i) {
    const int32_t x = inc[i];
    int64_t j = i - 1;
    while (j >= b && inc[j] > x) {
      inc[j + 1] = inc[j];
      --j;
    }
    inc[j + 1] = x;
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of int64 values: per-workgroup scans, the workgroup totals
// scanned on the host (a few thousand values), then the offsets added.

__global__ __launch_bounds__(256) void k_scan_local(const int64_t* in, int64_t* out, int64_t n,
                                                    int64_t* block_sum) {
  __shared__ int64_t part[256];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanBlock;
  const int t = threadIdx.x;
  int64_t v[4], s = 0;
  for (int k = 0; k < 4; ++k) {
    const int64_t i = base + 4 * t + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {  // inclusive Hillis-Steele scan of the parts
    const int64_t add = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int64_t run = part[t] - s;  // exclusive prefix of this thread's four values
  for (int k = 0; k < 4; ++k) {
    const int64_t i = base + 4 * t + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (t == 255) block_sum[blockIdx.x] = part[255];
}

__global__ void k_scan_add(int64_t* out, int64_t n, const int64_t* block_off) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] += block_off[i / kScanBlock];
}

// out[0..n) = exclusive scan of in[0..n); returns the total.
int64_t scan_exclusive(const int64_t* in, int64_t* out, int64_t n) {
  if (n == 0) return 0;
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  Buf bs(sizeof(int64_t) * nb);
  hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(256), 0, nullptr, in, out, n, bs.as<int64_t>());
  ASM_HIP(hipGetLastError());
  std::vector<int64_t> h(nb);
  ASM_HIP(hipMemcpy(h.data(), bs.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost));
  int64_t run = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t x = h[b];
    h[b] = run;
    run += x;
  }
  ASM_HIP(hipMemcpy(bs.p, h.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256), dim3(256), 0, nullptr, out, n,
                     bs.as<int64_t>());
  ASM_HIP(hipGetLastError());
  return run;
}

// ---------------------------------------------------------------------------
// Node rows. FILL = false: per (node, DOF) the K and C row lengths and whether
// the DOF is kept. FILL = true: columns and values of the kept rows.
// err: bit 0 = a node has more than kNbMax neighbours, bit 1 = an entry in a
// column whose DOF was removed (not expected: the element matrices are
// symmetric, so a nonzero entry makes its column's row nonempty too).
struct RowOut {
  int64_t* kcnt = nullptr;  // n_nodes*max_df
  int64_t* ccnt = nullptr;
  int64_t* keep = nullptr;  // pass A writes it, pass B reads it
  const int64_t* newidx = nullptr;  // exclusive scan of keep
  const int64_t* kpos = nullptr;    // exclusive scan of kcnt (kept rows only)
  const int64_t* cpos = nullptr;
  int32_t* kcol = nullptr;
  double* kval = nullptr;
  int32_t* ccol = nullptr;
  double* cval = nullptr;
  double* mval = nullptr;
  int64_t* gl = nullptr;
  int32_t* err = nullptr;
};

template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_node_rows(DMesh m, RowOut o) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= m.n_nodes) return;
  const int md = m.max_df;
  const int64_t q0 = m.inc_ptr[v], q1 = m.inc_ptr[v + 1];
  // distinct vertices of the incident elements, ascending
  int32_t nb[kNbMax];
  int nn = 0;
  for (int64_t q = q0; q < q1; ++q) {
    const int64_t e = m.inc[q];
    const int ne = nodes_of(m.typ[e]);
    for (int k = 0; k < ne; ++k) {
      const int32_t w = m.els[e * m.nper + k];
      int p = 0;
      while (p < nn && nb[p] < w) ++p;
      if (p < nn && nb[p] == w) continue;
      if (nn == kNbMax) {
        atomicOr(o.err, 1);
        return;
      }
      for (int r = nn; r > p; --r) nb[r] = nb[r - 1];
      nb[p] = w;
      ++nn;
    }
  }
  if (nn == 0) {  // a node of no element: every DOF removed
    for (int d = 1; d <= md; ++d) {
      const int64_t g = v * md + (d - 1);
      if constexpr (!FILL) {
        o.keep[g] = 0;
        o.kcnt[g] = 0;
        o.ccnt[g] = 0;
      } else {
        o.gl[g] = 0;
      }
    }
    return;
  }
  int self = 0;
  while (nb[self] != v) ++self;
  double kacc[kNbMax * kMaxDf], cacc[kNbMax * kMaxDf];
  for (int d = 1; d <= md; ++d) {
    const int64_t g = v * md + (d - 1);
    for (int i = 0; i < nn * md; ++i) {
      kacc[i] = 0.0;
      cacc[i] = 0.0;
    }
    double mm = 0.0, cdg = 0.0;
    if (keep_dof(m.mask, d)) {
      for (int64_t q = q0; q < q1; ++q) {
        const int64_t e = m.inc[q];
        const int32_t t = m.typ[e];
        if (!type_has_dof(t, d)) continue;
        const int32_t* ev = m.els + e * m.nper;
        const int ne = nodes_of(t);
        int a = 0;
        while (a < ne && ev[a] != v) ++a;
        double me, ce;
        element_row(
            m, e, a, d,
            [&](int b, int d2, double kv, double cv) {
              const int32_t w = ev[b];
              int p = 0;
              while (nb[p] != w) ++p;
              kacc[p * md + d2 - 1] += kv;
              cacc[p * md + d2 - 1] += cv;
            },
            &me, &ce);
        mm += me;
        cdg += ce;
      }
      cacc[self * md + d - 1] += cdg;
    }
    int64_t kc = 0, cc = 0;
    for (int i = 0; i < nn * md; ++i) {
      kc += kacc[i] != 0.0;
      cc += cacc[i] != 0.0;
    }
    const bool kept = kc > 0 || cc > 0 || mm != 0.0;
    if constexpr (!FILL) {
      o.keep[g] = kept ? 1 : 0;
      o.kcnt[g] = kept ? kc : 0;
      o.ccnt[g] = kept ? cc : 0;
    } else {
      if (!kept) {
        o.gl[g] = 0;
        continue;
      }
      const int64_t r = o.newidx[g];
      o.gl[g] = r + 1;
      o.mval[r] = mm;
      int64_t pk = o.kpos[g], pc = o.cpos[g];
      for (int p = 0; p < nn; ++p)
        for (int d2 = 1; d2 <= md; ++d2) {
          const int i = p * md + d2 - 1;
          if (kacc[i] == 0.0 && cacc[i] == 0.0) continue;
          const int64_t gc = static_cast<int64_t>(nb[p]) * md + (d2 - 1);
          if (!o.keep[gc]) {  // column of a removed DOF
            atomicOr(o.err, 2);
            continue;
          }
          const int64_t c = o.newidx[gc];
          if (kacc[i] != 0.0) {
            o.kcol[pk] = static_cast<int32_t>(c);
            o.kval[pk++] = kacc[i];
          }
          if (cacc[i] != 0.0) {
            o.ccol[pc] = static_cast<int32_t>(c);
            o.cval[pc++] = cacc[i];
          }
        }
    }
  }
}

// Compacted row pointers: rowptr[newidx[g]] = pos[g] for kept g.
__global__ void k_rowptr(int64_t n, const int64_t* keep, const int64_t* newidx,
                         const int64_t* pos, int64_t* rowptr) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g < n && keep[g]) rowptr[newidx[g]] = pos[g];
}

// ---------------------------------------------------------------------------
// Element templates of the element-by-element operator: elements of one type
// whose node offsets from vertex 1, rounded to 2^-24 of the largest element
// extent, and material agree share a template. key = sum_i q_i mult_i over
// the rounded offsets and the material (wrapping 64-bit arithmetic); distinct
// keys go into a small open-addressing set, each slot keeping its smallest
// element index (the representative).

constexpr int kTmplSetSize = 4096;  // slots of the key set (power of two)
constexpr uint64_t kTmplEmpty = ~0ull;

struct TmplMesh {
  int64_t n_el = 0;
  int32_t npe = 0, dim = 0;
  const double* nds = nullptr;
  const int32_t* els = nullptr;  // n_el x npe, 0-based
  const int32_t* mat = nullptr;
};

__global__ void k_tmpl_extent(TmplMesh m, unsigned long long* ext) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double x = 0.0;
  if (e < m.n_el) {
    const int32_t* v = m.els + e * m.npe;
    for (int b = 1; b < m.npe; ++b)
      for (int c = 0; c < m.dim; ++c)
        x = fmax(x, fabs(m.nds[static_cast<int64_t>(v[b]) * m.dim + c] -
                         m.nds[static_cast<int64_t>(v[0]) * m.dim + c]));
  }
  // non-negative doubles order as their bit patterns
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_down(x, off));
  if ((threadIdx.x & 63) == 0)
    atomicMax(ext, static_cast<unsigned long long>(__double_as_longlong(x)));
}

__device__ __forceinline__ uint64_t tmpl_key(const TmplMesh& m, int64_t e, double scale,
                                             const uint64_t* mult) {
  const int32_t* v = m.els + e * m.npe;
  uint64_t key = 0;
  int i = 0;
  for (int b = 1; b < m.npe; ++b)
    for (int c = 0; c < m.dim; ++c, ++i) {
      const double d = m.nds[static_cast<int64_t>(v[b]) * m.dim + c] -
                       m.nds[static_cast<int64_t>(v[0]) * m.dim + c];
      key += static_cast<uint64_t>(static_cast<int64_t>(rint(d * scale))) * mult[i];
    }
  key += static_cast<uint64_t>(static_cast<int64_t>(m.mat[e])) * mult[i];
  return key;
}

// Inserts every element's key; err bit 0: more than max_keys distinct keys,
// bit 1: a key equal to the empty marker. The lanes of a wave share their
// keys first: one lane per distinct key (the lowest, whose element is the
// smallest) inserts it, so a wave of one template costs one probe sequence
// and one atomicMin instead of 64 on the same slot.
__global__ void k_tmpl_insert(TmplMesh m, double scale, const uint64_t* mult,
                              unsigned long long* set_key, unsigned long long* set_first,
                              int32_t* el_slot, int32_t* n_keys, int32_t max_keys, int32_t* err) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int lane = static_cast<int>(threadIdx.x & 63);
  bool pending = e < m.n_el && !(*reinterpret_cast<volatile int32_t*>(err) & 1);
  const uint64_t key = pending ? tmpl_key(m, e, scale, mult) : 0ull;
  if (pending && key == kTmplEmpty) {
    atomicOr(err, 2);
    pending = false;
  }
  for (;;) {
    const unsigned long long todo = __ballot(pending);
    if (todo == 0ull) break;
    const int leader = __ffsll(static_cast<long long>(todo)) - 1;
    const uint64_t lk = __shfl(key, leader);
    int32_t slot = -1;
    if (lane == leader) {
      uint32_t h = static_cast<uint32_t>((lk ^ (lk >> 29)) * 0x9E3779B97F4A7C15ull >> 52) &
                   (kTmplSetSize - 1);
      for (int probe = 0; probe < kTmplSetSize; ++probe) {
        const unsigned long long prev = atomicCAS(set_key + h, kTmplEmpty, lk);
        if (prev == kTmplEmpty && atomicAdd(n_keys, 1) >= max_keys) break;  // too many
        if (prev == kTmplEmpty || prev == lk) {
          atomicMin(set_first + h, static_cast<unsigned long long>(e));
          slot = static_cast<int32_t>(h);
          break;
        }
        h = (h + 1) & (kTmplSetSize - 1);
      }
      if (slot < 0) atomicOr(err, 1);
    }
    slot = __shfl(slot, leader);
    if (pending && key == lk) {
      if (slot >= 0) el_slot[e] = slot;
      pending = false;
    }
  }
}

// Template index of every element and, for triangles, J_rep / J_e with the
// reference's absolute-coordinate J (fn_el_CPE3.m:56); err bit 2: an element
// whose rounded offsets differ from its representative's (a hash collision).
#pragma clang fp contract(off)
__global__ void k_tmpl_assign(TmplMesh m, double scale, const int32_t* slot_tmpl,
                              const int64_t* first, const int32_t* el_slot, int32_t* el_tmpl,
                              double* jfac, int32_t* err) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m.n_el) return;
  const int32_t t = slot_tmpl[el_slot[e]];
  el_tmpl[e] = t;
  const int64_t r = first[t];
  const int32_t* v = m.els + e * m.npe;
  const int32_t* w = m.els + r * m.npe;
  auto X = [&](const int32_t* p, int b, int c) {
    return m.nds[static_cast<int64_t>(p[b]) * m.dim + c];
  };
  bool same = m.mat[e] == m.mat[r];
  for (int b = 1; b < m.npe; ++b)
    for (int c = 0; c < m.dim; ++c)
      same = same && rint((X(v, b, c) - X(v, 0, c)) * scale) == rint((X(w, b, c) - X(w, 0, c)) * scale);
  if (!same) atomicOr(err, 4);
  if (jfac != nullptr) {
    auto J = [&](const int32_t* p) {
      const double x1 = X(p, 0, 0), y1 = X(p, 0, 1), x2 = X(p, 1, 0), y2 = X(p, 1, 1),
                   x3 = X(p, 2, 0), y3 = X(p, 2, 1);
      return x1 * y2 - y1 * x2 - x1 * y3 + y1 * x3 + x2 * y3 - y2 * x3;
    };
    jfac[e] = m.npe == 3 ? J(w) / J(v) : 1.0;
  }
}
#pragma clang fp contract(on)

struct DevAssembly {
  int32_t device = 0;
  int64_t n_nodes = 0, n_dof = 0, nnz_k = 0, nnz_c = 0;
  int32_t max_df = 0;
  Buf krow, kcol, kval, crow, ccol, cval, mval, gl;
};

inline unsigned grid(int64_t n, int threads) {
  return static_cast<unsigned>((n + threads - 1) / threads);
}

}  // namespace

extern "C" {

bfe_status bfe_assemble_device(int32_t device, int64_t n_nodes, int32_t dim, const double* nds,
                               int64_t n_el, int32_t nper, const int32_t* els,
                               const int32_t* el_type, const int32_t* el_mat, const double* el_abs,
                               int32_t n_mat, const double* mat_D, const double* mat_rho,
                               int32_t dof_mask, double damping_power_law, double max_damping,
                               double max_stiffness_reduction, int64_t* n_dof, int64_t* nnz_k,
                               int64_t* nnz_c, int32_t* max_df, void** work) {
  try {
    if (!n_dof || !nnz_k || !nnz_c || !max_df || !work) throw std::invalid_argument("NULL output");
    *work = nullptr;
    if (n_nodes <= 0 || n_nodes > INT32_MAX || n_el <= 0 || n_el > INT32_MAX ||
        (dim != 2 && dim != 3) || nper < 2)
      throw std::invalid_argument("bad mesh sizes");
    if (!nds || !els || !el_type || !el_mat || !mat_D || !mat_rho)
      throw std::invalid_argument("NULL mesh array");
    if (!(max_stiffness_reduction > 0.0)) throw std::invalid_argument("max_stiffness_reduction <= 0");
    // element checks and max_df (fn_find_dof_in_use_and_max_dof_per_el.m)
    bool seen[4] = {false, false, false, false};
    for (int64_t e = 0; e < n_el; ++e) {
      const int32_t t = el_type[e];
      if (t < 0 || t > 3) throw std::invalid_argument("unsupported element type");
      if (nodes_of(t) > nper) throw std::invalid_argument("element type needs more node columns");
      if (t == kC3D4 && dim != 3) throw std::invalid_argument("C3D4 needs 3D nodes");
      if (t != kASI2D2 && (el_mat[e] < 0 || el_mat[e] >= n_mat))
        throw std::invalid_argument("material out of range");
      seen[t] = true;
      for (int k = 0; k < nodes_of(t); ++k) {
        const int32_t v = els[e * nper + k];
        if (v < 0 || v >= n_nodes) throw std::invalid_argument("node index out of range");
      }
    }
    int dmax = 0;
    for (int t = 0; t < 4; ++t) {
      if (!seen[t]) continue;
      for (int d = 1; d <= 4; ++d) {
        const bool has = t == kAC2D3 ? d == 4 : (t == kASI2D2 ? true : d <= 3);
        if (has && keep_dof(dof_mask, d)) dmax = std::max(dmax, d);
      }
    }
    if (dmax == 0) throw std::invalid_argument("no DOF in use");
    ASM_HIP(hipSetDevice(device));
    // materials with a zero row for interface elements (el_mat -1 allowed only there)
    std::vector<int32_t> mat(el_mat, el_mat + n_el);
    for (int64_t e = 0; e < n_el; ++e)
      if (el_type[e] == kASI2D2) mat[e] = 0;
    std::vector<double> Dh(mat_D, mat_D + static_cast<size_t>(std::max(n_mat, 1)) * 36);
    std::vector<double> rh(mat_rho, mat_rho + std::max(n_mat, 1));
    if (n_mat == 0) {
      Dh.assign(36, 0.0);
      rh.assign(1, 0.0);
    }
    Buf d_nds = upload(nds, static_cast<size_t>(n_nodes) * dim);
    Buf d_els = upload(els, static_cast<size_t>(n_el) * nper);
    Buf d_typ = upload(el_type, static_cast<size_t>(n_el));
    Buf d_mat = upload(mat.data(), mat.size());
    Buf d_abs;
    if (el_abs) d_abs = upload(el_abs, static_cast<size_t>(n_el));
    Buf d_D = upload(Dh.data(), Dh.size());
    Buf d_rho = upload(rh.data(), rh.size());
    DMesh m;
    m.n_nodes = n_nodes;
    m.n_el = n_el;
    m.dim = dim;
    m.nper = nper;
    m.mask = dof_mask;
    m.max_df = dmax;
    m.p_law = damping_power_law;
    m.max_damp = max_damping;
    m.log_msr = std::log(max_stiffness_reduction);
    m.nds = d_nds.as<double>();
    m.els = d_els.as<int32_t>();
    m.typ = d_typ.as<int32_t>();
    m.mat = d_mat.as<int32_t>();
    m.abs_i = el_abs ? d_abs.as<double>() : nullptr;
    m.D = d_D.as<double>();
    m.rho = d_rho.as<double>();
    // 1. incidence
    Buf cnt(sizeof(int64_t) * (n_nodes + 1));
    ASM_HIP(hipMemset(cnt.p, 0, cnt.bytes));
    hipLaunchKernelGGL(k_inc_count, dim3(grid(n_el, 256)), dim3(256), 0, nullptr, m,
                       cnt.as<int64_t>());
    ASM_HIP(hipGetLastError());
    Buf ptr(sizeof(int64_t) * (n_nodes + 1));
    const int64_t n_inc = scan_exclusive(cnt.as<int64_t>(), ptr.as<int64_t>(), n_nodes + 1);
    cnt.release();
    Buf cursor(sizeof(int32_t) * n_nodes);
    ASM_HIP(hipMemset(cursor.p, 0, cursor.bytes));
    Buf inc(sizeof(int32_t) * std::max<int64_t>(n_inc, 1));
    hipLaunchKernelGGL(k_inc_fill, dim3(grid(n_el, 256)), dim3(256), 0, nullptr, m,
                       ptr.as<int64_t>(), cursor.as<int32_t>(), inc.as<int32_t>());
    ASM_HIP(hipGetLastError());
    cursor.release();
    hipLaunchKernelGGL(k_inc_sort, dim3(grid(n_nodes, 256)), dim3(256), 0, nullptr, n_nodes,
                       ptr.as<int64_t>(), inc.as<int32_t>());
    ASM_HIP(hipGetLastError());
    m.inc_ptr = ptr.as<int64_t>();
    m.inc = inc.as<int32_t>();
    // 2. pass A
    const int64_t ng = n_nodes * dmax;
    Buf keep(sizeof(int64_t) * ng), kcnt(sizeof(int64_t) * ng), ccnt(sizeof(int64_t) * ng);
    Buf err(sizeof(int32_t));
    ASM_HIP(hipMemset(err.p, 0, err.bytes));
    RowOut o;
    o.keep = keep.as<int64_t>();
    o.kcnt = kcnt.as<int64_t>();
    o.ccnt = ccnt.as<int64_t>();
    o.err = err.as<int32_t>();
    hipLaunchKernelGGL((k_node_rows<false>), dim3(grid(n_nodes, kThreads)), dim3(kThreads), 0,
                       nullptr, m, o);
    ASM_HIP(hipGetLastError());
    int32_t herr = 0;
    ASM_HIP(hipMemcpy(&herr, err.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (herr & 1)
      throw std::invalid_argument("a node has more than " + std::to_string(kNbMax) +
                                  " neighbouring nodes");
    // 3. numbering and row pointers
    Buf newidx(sizeof(int64_t) * ng), kpos(sizeof(int64_t) * ng), cpos(sizeof(int64_t) * ng);
    const int64_t nd = scan_exclusive(keep.as<int64_t>(), newidx.as<int64_t>(), ng);
    const int64_t zk = scan_exclusive(kcnt.as<int64_t>(), kpos.as<int64_t>(), ng);
    const int64_t zc = scan_exclusive(ccnt.as<int64_t>(), cpos.as<int64_t>(), ng);
    kcnt.release();
    ccnt.release();
    if (nd > INT32_MAX) throw std::invalid_argument("more than 2^31 DOF");
    auto W = std::make_unique<DevAssembly>();
    W->device = device;
    W->n_nodes = n_nodes;
    W->n_dof = nd;
    W->nnz_k = zk;
    W->nnz_c = zc;
    W->max_df = dmax;
    W->krow = Buf(sizeof(int64_t) * (nd + 1));
    W->crow = Buf(sizeof(int64_t) * (nd + 1));
    hipLaunchKernelGGL(k_rowptr, dim3(grid(ng, 256)), dim3(256), 0, nullptr, ng,
                       keep.as<int64_t>(), newidx.as<int64_t>(), kpos.as<int64_t>(),
                       W->krow.as<int64_t>());
    hipLaunchKernelGGL(k_rowptr, dim3(grid(ng, 256)), dim3(256), 0, nullptr, ng,
                       keep.as<int64_t>(), newidx.as<int64_t>(), cpos.as<int64_t>(),
                       W->crow.as<int64_t>());
    ASM_HIP(hipGetLastError());
    ASM_HIP(hipMemcpy(W->krow.as<int64_t>() + nd, &zk, sizeof(int64_t), hipMemcpyHostToDevice));
    ASM_HIP(hipMemcpy(W->crow.as<int64_t>() + nd, &zc, sizeof(int64_t), hipMemcpyHostToDevice));
    // 4. pass B
    W->kcol = Buf(sizeof(int32_t) * std::max<int64_t>(zk, 1));
    W->kval = Buf(sizeof(double) * std::max<int64_t>(zk, 1));
    W->ccol = Buf(sizeof(int32_t) * std::max<int64_t>(zc, 1));
    W->cval = Buf(sizeof(double) * std::max<int64_t>(zc, 1));
    W->mval = Buf(sizeof(double) * std::max<int64_t>(nd, 1));
    W->gl = Buf(sizeof(int64_t) * ng);
    o.newidx = newidx.as<int64_t>();
    o.kpos = kpos.as<int64_t>();
    o.cpos = cpos.as<int64_t>();
    o.kcol = W->kcol.as<int32_t>();
    o.kval = W->kval.as<double>();
    o.ccol = W->ccol.as<int32_t>();
    o.cval = W->cval.as<double>();
    o.mval = W->mval.as<double>();
    o.gl = W->gl.as<int64_t>();
    hipLaunchKernelGGL((k_node_rows<true>), dim3(grid(n_nodes, kThreads)), dim3(kThreads), 0,
                       nullptr, m, o);
    ASM_HIP(hipGetLastError());
    ASM_HIP(hipMemcpy(&herr, err.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (herr & 2) throw std::invalid_argument("an entry in a removed column (asymmetric element)");
    *n_dof = nd;
    *nnz_k = zk;
    *nnz_c = zc;
    *max_df = dmax;
    *work = W.release();
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    bfe::set_assembly_error(std::string("invalid argument: ") + e.what());
    return BFE_ERR_ARG;
  } catch (const dev_failure& e) {
    (void)hipGetLastError();
    bfe::set_assembly_error(std::string("HIP error: ") + e.what());
    return BFE_ERR_HIP;
  } catch (const std::bad_alloc&) {
    bfe::set_assembly_error("out of memory");
    return BFE_ERR_NOMEM;
  } catch (const std::exception& e) {
    bfe::set_assembly_error(std::string("internal error: ") + e.what());
    return BFE_ERR_INTERNAL;
  }
}

bfe_status bfe_assemble_device_fill(void* work, int64_t* k_rowptr, int32_t* k_col, double* k_val,
                                    int64_t* c_rowptr, int32_t* c_col, double* c_val, double* m,
                                    int64_t* gl_lookup) {
  std::unique_ptr<DevAssembly> W(static_cast<DevAssembly*>(work));
  try {
    if (!W) throw std::invalid_argument("NULL work object");
    if (!k_rowptr || !k_col || !k_val || !c_rowptr || !c_col || !c_val || !m || !gl_lookup)
      throw std::invalid_argument("NULL output");
    ASM_HIP(hipSetDevice(W->device));
    auto down = [](void* dst, const Buf& src, size_t bytes) {
      if (bytes) ASM_HIP(hipMemcpy(dst, src.p, bytes, hipMemcpyDeviceToHost));
    };
    down(k_rowptr, W->krow, sizeof(int64_t) * (W->n_dof + 1));
    down(c_rowptr, W->crow, sizeof(int64_t) * (W->n_dof + 1));
    down(k_col, W->kcol, sizeof(int32_t) * W->nnz_k);
    down(k_val, W->kval, sizeof(double) * W->nnz_k);
    down(c_col, W->ccol, sizeof(int32_t) * W->nnz_c);
    down(c_val, W->cval, sizeof(double) * W->nnz_c);
    down(m, W->mval, sizeof(double) * W->n_dof);
    down(gl_lookup, W->gl, sizeof(int64_t) * W->n_nodes * W->max_df);
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    bfe::set_assembly_error(std::string("invalid argument: ") + e.what());
    return BFE_ERR_ARG;
  } catch (const dev_failure& e) {
    (void)hipGetLastError();
    bfe::set_assembly_error(std::string("HIP error: ") + e.what());
    return BFE_ERR_HIP;
  } catch (const std::exception& e) {
    bfe::set_assembly_error(std::string("internal error: ") + e.what());
    return BFE_ERR_INTERNAL;
  }
}

bfe_status bfe_assemble_device_free(void* work) {
  delete static_cast<DevAssembly*>(work);
  return BFE_OK;
}

bfe_status bfe_element_templates(int32_t device, int64_t n_nodes, int32_t dim, const double* nds,
                                 int64_t n_el, int32_t npe, const int32_t* els,
                                 const int32_t* el_mat, const int64_t* mult, int32_t max_templates,
                                 int32_t* el_tmpl, int64_t* first, int32_t* n_tmpl,
                                 double* jfactor) {
  try {
    if (!n_tmpl) throw std::invalid_argument("NULL output");
    *n_tmpl = 0;
    if (n_nodes <= 0 || n_el <= 0 || n_el > INT32_MAX || (dim != 2 && dim != 3) || npe < 2 ||
        npe > 8 || max_templates < 1 || max_templates > kTmplSetSize / 4)
      throw std::invalid_argument("bad sizes");
    if (!nds || !els || !el_mat || !mult || !el_tmpl || !first)
      throw std::invalid_argument("NULL array");
    if (jfactor != nullptr && (npe != 3 || dim != 2))
      throw std::invalid_argument("the Jacobian ratio is defined for 2D triangles");
    for (int64_t i = 0; i < n_el * npe; ++i)
      if (els[i] < 0 || els[i] >= n_nodes) throw std::invalid_argument("node index out of range");
    ASM_HIP(hipSetDevice(device));
    const int nm = (npe - 1) * dim + 1;
    Buf d_nds = upload(nds, static_cast<size_t>(n_nodes) * dim);
    Buf d_els = upload(els, static_cast<size_t>(n_el) * npe);
    Buf d_mat = upload(el_mat, static_cast<size_t>(n_el));
    Buf d_mult = upload(mult, static_cast<size_t>(nm));
    TmplMesh m;
    m.n_el = n_el;
    m.npe = npe;
    m.dim = dim;
    m.nds = d_nds.as<double>();
    m.els = d_els.as<int32_t>();
    m.mat = d_mat.as<int32_t>();
    Buf ext(sizeof(unsigned long long));
    ASM_HIP(hipMemset(ext.p, 0, ext.bytes));
    hipLaunchKernelGGL(k_tmpl_extent, dim3(grid(n_el, 256)), dim3(256), 0, nullptr, m,
                       ext.as<unsigned long long>());
    ASM_HIP(hipGetLastError());
    unsigned long long eb = 0;
    ASM_HIP(hipMemcpy(&eb, ext.p, sizeof(eb), hipMemcpyDeviceToHost));
    double e_max;
    std::memcpy(&e_max, &eb, sizeof(e_max));
    if (!(e_max > 0.0)) throw std::invalid_argument("elements of zero extent");
    const double scale = 16777216.0 / e_max;  // 2^24 / extent
    Buf set_key(sizeof(unsigned long long) * kTmplSetSize);
    Buf set_first(sizeof(unsigned long long) * kTmplSetSize);
    ASM_HIP(hipMemset(set_key.p, 0xff, set_key.bytes));
    ASM_HIP(hipMemset(set_first.p, 0xff, set_first.bytes));
    Buf el_slot(sizeof(int32_t) * n_el), cnt(sizeof(int32_t)), err(sizeof(int32_t));
    ASM_HIP(hipMemset(cnt.p, 0, cnt.bytes));
    ASM_HIP(hipMemset(err.p, 0, err.bytes));
    hipLaunchKernelGGL(k_tmpl_insert, dim3(grid(n_el, 256)), dim3(256), 0, nullptr, m, scale,
                       d_mult.as<uint64_t>(), set_key.as<unsigned long long>(),
                       set_first.as<unsigned long long>(), el_slot.as<int32_t>(),
                       cnt.as<int32_t>(), max_templates, err.as<int32_t>());
    ASM_HIP(hipGetLastError());
    int32_t herr = 0;
    ASM_HIP(hipMemcpy(&herr, err.p, sizeof(herr), hipMemcpyDeviceToHost));
    if (herr & 1) return BFE_OK;  // more than max_templates shapes: *n_tmpl stays 0
    if (herr & 2) throw std::invalid_argument("an element key equals the empty marker");
    std::vector<unsigned long long> hk(kTmplSetSize), hf(kTmplSetSize);
    ASM_HIP(hipMemcpy(hk.data(), set_key.p, set_key.bytes, hipMemcpyDeviceToHost));
    ASM_HIP(hipMemcpy(hf.data(), set_first.p, set_first.bytes, hipMemcpyDeviceToHost));
    // templates in ascending order of the key read as a signed integer
    std::vector<int32_t> used;
    for (int32_t h = 0; h < kTmplSetSize; ++h)
      if (hk[h] != kTmplEmpty) used.push_back(h);
    std::sort(used.begin(), used.end(), [&](int32_t a, int32_t b) {
      return static_cast<int64_t>(hk[a]) < static_cast<int64_t>(hk[b]);
    });
    std::vector<int32_t> slot_tmpl(kTmplSetSize, -1);
    std::vector<int64_t> fst(used.size());
    for (size_t t = 0; t < used.size(); ++t) {
      slot_tmpl[used[t]] = static_cast<int32_t>(t);
      fst[t] = static_cast<int64_t>(hf[used[t]]);
    }
    Buf d_st = upload(slot_tmpl.data(), slot_tmpl.size());
    Buf d_first = upload(fst.data(), fst.size());
    Buf d_tmpl(sizeof(int32_t) * n_el);
    Buf d_jf;
    if (jfactor) d_jf = Buf(sizeof(double) * n_el);
    hipLaunchKernelGGL(k_tmpl_assign, dim3(grid(n_el, 256)), dim3(256), 0, nullptr, m, scale,
                       d_st.as<int32_t>(), d_first.as<int64_t>(), el_slot.as<int32_t>(),
                       d_tmpl.as<int32_t>(), jfactor ? d_jf.as<double>() : nullptr,
                       err.as<int32_t>());
    ASM_HIP(hipGetLastError());
    ASM_HIP(hipMemcpy(&herr, err.p, sizeof(herr), hipMemcpyDeviceToHost));
    if (herr & 4) throw std::invalid_argument("two element shapes share a key (hash collision)");
    ASM_HIP(hipMemcpy(el_tmpl, d_tmpl.p, sizeof(int32_t) * n_el, hipMemcpyDeviceToHost));
    if (jfactor)
      ASM_HIP(hipMemcpy(jfactor, d_jf.p, sizeof(double) * n_el, hipMemcpyDeviceToHost));
    std::memcpy(first, fst.data(), sizeof(int64_t) * fst.size());
    *n_tmpl = static_cast<int32_t>(fst.size());
    return BFE_OK;
  } catch (const std::invalid_argument& e) {
    bfe::set_assembly_error(std::string("invalid argument: ") + e.what());
    return BFE_ERR_ARG;
  } catch (const dev_failure& e) {
    (void)hipGetLastError();
    bfe::set_assembly_error(std::string("HIP error: ") + e.what());
    return BFE_ERR_HIP;
  } catch (const std::bad_alloc&) {
    bfe::set_assembly_error("out of memory");
    return BFE_ERR_NOMEM;
  } catch (const std::exception& e) {
    bfe::set_assembly_error(std::string("internal error: ") + e.what());
    return BFE_ERR_INTERNAL;
  }
}

}  // extern "C"

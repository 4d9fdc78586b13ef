// Per-gate state-vector kernels (the plain, one-launch-per-gate path).
//
// Replaces numpybk.tensormul (quantumflow/backend/numpybk.py:159-214) as
// called from Gate.run (quantumflow/ops.py:161-166): out = G applied to the
// axes `targets` of in.  Every kernel assigns one amplitude group (the 2^k
// amplitudes that differ only in the target bits) to one work item, so the
// 16-byte complex128 loads of consecutive lanes are consecutive in memory
// whenever the lowest index bits are not targets, and out == in is safe.
// Amplitudes are streamed with nontemporal loads and stores (each is
// touched once per gate).
#include <cmath>
#include <cstring>

#include "qfx_common.h"

namespace {

constexpr int kBlock = 256;

template <int K>
struct DenseArg {
  double2 m[1 << (2 * K)];   // row-major, row = output
  uint64_t off[1 << K];      // address offset of gate-local index r
  int sorted[K];             // target bits ascending
};

// groups per thread: keep ~8 amplitude loads in flight per lane
template <int K>
constexpr int groups_per_thread() { return K == 1 ? 4 : (K == 2 ? 2 : 1); }

template <int K, int G>
__global__ __launch_bounds__(kBlock) void dense_kernel(double2* out, const double2* in,
                                                      uint64_t ngroups, DenseArg<K> a) {
  constexpr int D = 1 << K;
  const uint64_t g0 = (uint64_t)blockIdx.x * (kBlock * G) + threadIdx.x;
  double2 v[G][D];
  uint64_t base[G];
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const uint64_t g = g0 + (uint64_t)u * kBlock;
    base[u] = insert_zeros<K>(g, a.sorted);
    if (g < ngroups) {
#pragma unroll
      for (int r = 0; r < D; ++r) v[u][r] = ld_stream(in + base[u] + a.off[r]);
    }
  }
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const uint64_t g = g0 + (uint64_t)u * kBlock;
    if (g < ngroups) {
#pragma unroll
      for (int r = 0; r < D; ++r) {
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < D; ++c) cmac(acc, a.m[r * D + c], v[u][c]);
        st_stream(out + base[u] + a.off[r], acc);
      }
    }
  }
}

// Rows of a diagonal or monomial gate: out row r <- phase[r] * in row src[r].
struct SparseArg {
  uint64_t off[32];     // address offset of output row r
  uint64_t srcoff[32];  // address offset of the source row of r
  double2 ph[32];       // phase of row r
  uint32_t active;      // rows written (all rows out of place)
  uint32_t has_phase;   // rows whose phase is not exactly 1
  int sorted[5];
};

template <int K, int G>
__global__ __launch_bounds__(kBlock) void sparse_kernel(double2* out, const double2* in,
                                                       uint64_t ngroups, SparseArg a) {
  constexpr int D = 1 << K;
  const uint64_t g0 = (uint64_t)blockIdx.x * (kBlock * G) + threadIdx.x;
  double2 v[G][D];
  uint64_t base[G];
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const uint64_t g = g0 + (uint64_t)u * kBlock;
    base[u] = insert_zeros<K>(g, a.sorted);
    if (g < ngroups) {
#pragma unroll
      for (int r = 0; r < D; ++r)
        if (a.active & (1u << r)) v[u][r] = ld_stream(in + base[u] + a.srcoff[r]);
    }
  }
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const uint64_t g = g0 + (uint64_t)u * kBlock;
    if (g < ngroups) {
#pragma unroll
      for (int r = 0; r < D; ++r) {
        if (a.active & (1u << r)) {
          const double2 x = (a.has_phase & (1u << r)) ? cmul(a.ph[r], v[u][r]) : v[u][r];
          st_stream(out + base[u] + a.off[r], x);
        }
      }
    }
  }
}

// General dense gate with the matrix in device memory; one output amplitude
// per lane.  Needs out != in.  The 2^k column offsets are built in LDS.
struct DevTargets {
  int t[QFX_MAX_DENSE_K];  // bit of gate qubit j (MSB first)
};

__global__ __launch_bounds__(kBlock) void dense_dev_kernel(double2* __restrict__ out,
                                                          const double2* __restrict__ in,
                                                          uint64_t len, int k, DevTargets tg,
                                                          const double2* __restrict__ m) {
  extern __shared__ uint64_t s_off[];
  const int D = 1 << k;
  uint64_t tmask = 0;
  for (int j = 0; j < k; ++j) tmask |= (uint64_t)1 << tg.t[j];
  for (int c = threadIdx.x; c < D; c += kBlock) {
    uint64_t off = 0;
    for (int j = 0; j < k; ++j)
      if ((c >> (k - 1 - j)) & 1) off |= (uint64_t)1 << tg.t[j];
    s_off[c] = off;
  }
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= len) return;
  unsigned r = 0;
  for (int j = 0; j < k; ++j) r = (r << 1) | (unsigned)((i >> tg.t[j]) & 1);
  const uint64_t base = i & ~tmask;
  double2 acc = make_double2(0.0, 0.0);
  const double2* row = m + (uint64_t)r * D;
  for (int c = 0; c < D; ++c) cmac(acc, row[c], in[base + s_off[c]]);
  out[i] = acc;
}

// out[i] = in[j] where bit src[b] of j equals bit b of i (a relabelling of
// the index bits; writes are coalesced, reads gather).
struct PermArg {
  int src[QFX_MAX_QUBITS];
};

__global__ __launch_bounds__(kBlock) void permute_bits_kernel(double2* __restrict__ out,
                                                             const double2* __restrict__ in,
                                                             uint64_t len, int n, PermArg a) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= len) return;
  uint64_t j = 0;
  for (int b = 0; b < n; ++b) j |= ((i >> b) & 1) << a.src[b];
  out[i] = in[j];
}

__global__ __launch_bounds__(kBlock) void fill_kernel(double2* psi, uint64_t len, double2 first,
                                                     double2 rest) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < len) psi[i] = (i == 0) ? first : rest;
}

template <int K>
int launch_dense(hipStream_t s, double2* out, const double2* in, int n, const int* targets,
                 const qfx_c128* matrix) {
  DenseArg<K> a;
  constexpr int D = 1 << K;
  for (int i = 0; i < D * D; ++i) a.m[i] = make_double2(matrix[i].re, matrix[i].im);
  for (unsigned r = 0; r < (unsigned)D; ++r) a.off[r] = qfx::local_offset(K, targets, r);
  qfx::sort_bits(K, targets, a.sorted);
  constexpr int G = groups_per_thread<K>();
  const uint64_t ngroups = (uint64_t)1 << (n - K);
  const uint64_t blocks = (ngroups + kBlock * G - 1) / (kBlock * G);
  hipLaunchKernelGGL((dense_kernel<K, G>), dim3((unsigned)blocks), dim3(kBlock), 0, s, out, in,
                     ngroups, a);
  return qfx::check_launch("dense_kernel");
}

template <int K>
int launch_sparse_k(hipStream_t s, double2* out, const double2* in, int n, const SparseArg& a) {
  constexpr int G = K == 1 ? 4 : ((K == 2 || K == 4) ? 2 : 1);
  const uint64_t ngroups = (uint64_t)1 << (n - K);
  const uint64_t blocks = (ngroups + kBlock * G - 1) / (kBlock * G);
  hipLaunchKernelGGL((sparse_kernel<K, G>), dim3((unsigned)blocks), dim3(kBlock), 0, s, out, in,
                     ngroups, a);
  return qfx::check_launch("sparse_kernel");
}

int launch_sparse(hipStream_t s, double2* out, const double2* in, int n, int k,
                  const SparseArg& a) {
  switch (k) {
    case 1: return launch_sparse_k<1>(s, out, in, n, a);
    case 2: return launch_sparse_k<2>(s, out, in, n, a);
    case 3: return launch_sparse_k<3>(s, out, in, n, a);
    case 4: return launch_sparse_k<4>(s, out, in, n, a);
    case 5: return launch_sparse_k<5>(s, out, in, n, a);
  }
  return qfx::set_error(QFX_ERANGE, "sparse gate width out of range");
}

bool is_one(const qfx_c128& z) { return z.re == 1.0 && z.im == 0.0; }

// Build SparseArg for a diagonal (src == NULL) or monomial gate.
int build_sparse(int n, int k, const int* targets, const int* src, const qfx_c128* ph,
                 bool inplace, SparseArg* a, const char* fn) {
  int rc = qfx::check_targets(n, k, targets, 5, fn);
  if (rc) return rc;
  const int D = 1 << k;
  std::memset(a, 0, sizeof(*a));
  uint32_t seen = 0;
  for (int r = 0; r < D; ++r) {
    const int sr = src ? src[r] : r;
    if (sr < 0 || sr >= D) return qfx::set_error(QFX_EINVAL, std::string(fn) + ": bad source row");
    if (seen & (1u << sr))
      return qfx::set_error(QFX_EINVAL, std::string(fn) + ": source rows are not a permutation");
    seen |= 1u << sr;
    a->off[r] = qfx::local_offset(k, targets, r);
    a->srcoff[r] = qfx::local_offset(k, targets, sr);
    const qfx_c128 z = ph ? ph[r] : qfx_c128{1.0, 0.0};
    a->ph[r] = make_double2(z.re, z.im);
    const bool one = is_one(z);
    if (!one) a->has_phase |= 1u << r;
    if (!inplace || !one || sr != r) a->active |= 1u << r;
  }
  qfx::sort_bits(k, targets, a->sorted);
  return QFX_OK;
}

}  // namespace

extern "C" {

int qfx_apply_matrix(void* stream, void* out, const void* in, int n, int k, const int* targets,
                     const qfx_c128* matrix) {
  int rc = qfx::check_targets(n, k, targets, 3, "qfx_apply_matrix");
  if (rc) return rc;
  if (!out || !in || !matrix) return qfx::set_error(QFX_EINVAL, "qfx_apply_matrix: null pointer");
  hipStream_t s = qfx::as_stream(stream);
  auto* o = static_cast<double2*>(out);
  auto* i = static_cast<const double2*>(in);
  switch (k) {
    case 1: return launch_dense<1>(s, o, i, n, targets, matrix);
    case 2: return launch_dense<2>(s, o, i, n, targets, matrix);
    case 3: return launch_dense<3>(s, o, i, n, targets, matrix);
  }
  return qfx::set_error(QFX_ERANGE, "qfx_apply_matrix: k > 3 needs qfx_apply_matrix_dev");
}

int qfx_apply_matrix_dev(void* stream, void* out, const void* in, int n, int k,
                         const int* targets, const void* matrix) {
  int rc = qfx::check_targets(n, k, targets, QFX_MAX_DENSE_K, "qfx_apply_matrix_dev");
  if (rc) return rc;
  if (!out || !in || !matrix)
    return qfx::set_error(QFX_EINVAL, "qfx_apply_matrix_dev: null pointer");
  if (out == in) return qfx::set_error(QFX_EINVAL, "qfx_apply_matrix_dev: out must not alias in");
  DevTargets tg;
  std::memset(&tg, 0, sizeof(tg));
  for (int j = 0; j < k; ++j) tg.t[j] = targets[j];
  const uint64_t len = (uint64_t)1 << n;
  const uint64_t blocks = (len + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(dense_dev_kernel, dim3((unsigned)blocks), dim3(kBlock),
                     sizeof(uint64_t) << k, qfx::as_stream(stream), static_cast<double2*>(out),
                     static_cast<const double2*>(in), len, k, tg,
                     static_cast<const double2*>(matrix));
  return qfx::check_launch("dense_dev_kernel");
}

int qfx_apply_diagonal(void* stream, void* out, const void* in, int n, int k, const int* targets,
                       const qfx_c128* diag) {
  if (!out || !in || !diag) return qfx::set_error(QFX_EINVAL, "qfx_apply_diagonal: null pointer");
  SparseArg a;
  int rc = build_sparse(n, k, targets, nullptr, diag, false, &a, "qfx_apply_diagonal");
  if (rc) return rc;
  return launch_sparse(qfx::as_stream(stream), static_cast<double2*>(out),
                       static_cast<const double2*>(in), n, k, a);
}

int qfx_apply_monomial(void* stream, void* out, const void* in, int n, int k, const int* targets,
                       const int* src, const qfx_c128* phase) {
  if (!out || !in || !src) return qfx::set_error(QFX_EINVAL, "qfx_apply_monomial: null pointer");
  SparseArg a;
  int rc = build_sparse(n, k, targets, src, phase, false, &a, "qfx_apply_monomial");
  if (rc) return rc;
  return launch_sparse(qfx::as_stream(stream), static_cast<double2*>(out),
                       static_cast<const double2*>(in), n, k, a);
}

int qfx_apply_diagonal_inplace(void* stream, void* state, int n, int k, const int* targets,
                               const qfx_c128* diag) {
  if (!state || !diag)
    return qfx::set_error(QFX_EINVAL, "qfx_apply_diagonal_inplace: null pointer");
  SparseArg a;
  int rc = build_sparse(n, k, targets, nullptr, diag, true, &a, "qfx_apply_diagonal_inplace");
  if (rc) return rc;
  if (!a.active) return QFX_OK;
  auto* p = static_cast<double2*>(state);
  return launch_sparse(qfx::as_stream(stream), p, p, n, k, a);
}

int qfx_apply_monomial_inplace(void* stream, void* state, int n, int k, const int* targets,
                               const int* src, const qfx_c128* phase) {
  if (!state || !src)
    return qfx::set_error(QFX_EINVAL, "qfx_apply_monomial_inplace: null pointer");
  SparseArg a;
  int rc = build_sparse(n, k, targets, src, phase, true, &a, "qfx_apply_monomial_inplace");
  if (rc) return rc;
  if (!a.active) return QFX_OK;
  auto* p = static_cast<double2*>(state);
  return launch_sparse(qfx::as_stream(stream), p, p, n, k, a);
}

int qfx_permute_qubits(void* stream, void* out, const void* in, int n, const int* src_bit) {
  if (!out || !in || !src_bit) return qfx::set_error(QFX_EINVAL, "qfx_permute_qubits: null pointer");
  if (out == in) return qfx::set_error(QFX_EINVAL, "qfx_permute_qubits: out must not alias in");
  if (n < 1 || n > QFX_MAX_QUBITS)
    return qfx::set_error(QFX_ERANGE, "qfx_permute_qubits: n out of range");
  PermArg a;
  std::memset(&a, 0, sizeof(a));
  uint64_t seen = 0;
  for (int b = 0; b < n; ++b) {
    if (src_bit[b] < 0 || src_bit[b] >= n || ((seen >> src_bit[b]) & 1))
      return qfx::set_error(QFX_EINVAL, "qfx_permute_qubits: not a permutation");
    seen |= (uint64_t)1 << src_bit[b];
    a.src[b] = src_bit[b];
  }
  const uint64_t len = (uint64_t)1 << n;
  hipLaunchKernelGGL(permute_bits_kernel, dim3((unsigned)((len + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, qfx::as_stream(stream), static_cast<double2*>(out),
                     static_cast<const double2*>(in), len, n, a);
  return qfx::check_launch("permute_bits_kernel");
}

int qfx_zero_state(void* stream, void* psi, int n) {
  if (!psi) return qfx::set_error(QFX_EINVAL, "qfx_zero_state: null pointer");
  if (n < 0 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, "qfx_zero_state: n out of range");
  const uint64_t len = (uint64_t)1 << n;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((len + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double2*>(psi), len,
                     make_double2(1.0, 0.0), make_double2(0.0, 0.0));
  return qfx::check_launch("fill_kernel");
}

int qfx_plus_state(void* stream, void* psi, int n) {
  if (!psi) return qfx::set_error(QFX_EINVAL, "qfx_plus_state: null pointer");
  if (n < 0 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, "qfx_plus_state: n out of range");
  const uint64_t len = (uint64_t)1 << n;
  const double a = std::ldexp(1.0, -n / 2) * ((n & 1) ? 0.70710678118654752440 : 1.0);
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((len + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double2*>(psi), len, make_double2(a, 0.0),
                     make_double2(a, 0.0));
  return qfx::check_launch("fill_kernel");
}

}  // extern "C"

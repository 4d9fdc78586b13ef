// Density-matrix kernels.  A density matrix of n qubits is stored as its
// 2^n x 2^n complex128 matrix, row-major: the vectorised 2n-qubit tensor
// whose first n axes are the row (ket) indices and whose last n axes are
// the column (bra) indices, as Density.tensor holds it in the reference
// (quantumflow/states.py:264-344, rank-2 QubitVector, qubits.py:39-121).
// Gates and Kraus operators act on it through the state-vector gate kernels
// (rows by U, columns by conj(U)); these kernels add what that leaves:
//   qfx_density_outer     rho = psi psi^H   (State.asdensity, states.py:162-165)
//   qfx_density_identity  rho = s I         (mixed_density, states.py:322-326)
//   qfx_density_diag      Re rho_ii         (Density.probabilities, states.py:302-305)
//   qfx_density_trace     sum rho_ii        (Density.trace, states.py:277-279)
//   qfx_density_axpy      y += w x          (the weighted sum of Kraus.evolve,
//                                            channels.py:74-80)
//   qfx_density_partial_trace  trace over the qubits outside keep_mask
//                                           (QubitVector.partial_trace,
//                                            qubits.py:201-229)
#include "qfx_common.h"

namespace {

constexpr int kBlock = 256;
constexpr int kTraceBlock = 1024;

int grid_for(uint64_t len) {
  uint64_t g = (len + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

__global__ __launch_bounds__(kBlock) void outer_kernel(double2* __restrict__ rho,
                                                      const double2* __restrict__ psi, int n) {
  const uint64_t len = (uint64_t)1 << (2 * n), mask = ((uint64_t)1 << n) - 1;
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < len;
       t += (uint64_t)gridDim.x * kBlock) {
    const double2 a = psi[t >> n], b = psi[t & mask];
    // a conj(b)
    rho[t] = double2{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
  }
}

__global__ __launch_bounds__(kBlock) void identity_kernel(double2* __restrict__ rho, int n,
                                                         double s) {
  const uint64_t len = (uint64_t)1 << (2 * n), mask = ((uint64_t)1 << n) - 1;
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < len;
       t += (uint64_t)gridDim.x * kBlock)
    rho[t] = double2{(t >> n) == (t & mask) ? s : 0.0, 0.0};
}

__global__ __launch_bounds__(kBlock) void diag_kernel(double* __restrict__ out,
                                                     const double2* __restrict__ rho, int n) {
  const uint64_t dim = (uint64_t)1 << n;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * kBlock)
    out[i] = rho[i * (dim + 1)].x;
}

// One workgroup, a fixed split and a fixed-order tree: bitwise reproducible.
__global__ __launch_bounds__(kTraceBlock) void trace_kernel(double2* __restrict__ out,
                                                           const double2* __restrict__ rho,
                                                           int n) {
  const uint64_t dim = (uint64_t)1 << n;
  double re = 0.0, im = 0.0;
  for (uint64_t i = threadIdx.x; i < dim; i += kTraceBlock) {
    const double2 d = rho[i * (dim + 1)];
    re += d.x;
    im += d.y;
  }
  re = wave_sum(re);
  im = wave_sum(im);
  __shared__ double red[2][kTraceBlock / 64];
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = re;
    red[1][threadIdx.x >> 6] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0, m = 0.0;
    for (int w = 0; w < kTraceBlock / 64; ++w) {
      r += red[0][w];
      m += red[1][w];
    }
    out[0] = double2{r, m};
  }
}

__global__ __launch_bounds__(kBlock) void axpy_kernel(double2* __restrict__ y,
                                                     const double2* __restrict__ x, double w,
                                                     uint64_t len) {
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < len;
       t += (uint64_t)gridDim.x * kBlock) {
    const double2 a = x[t], b = y[t];
    y[t] = double2{fma(w, a.x, b.x), fma(w, a.y, b.y)};
  }
}

// out[r', c'] = sum_u rho[r, c], r = deposit(r', keep) | d(u), c =
// deposit(c', keep) | d(u), d(u) = the u-th index over the traced bits
// (stepped by one carry-add).  One thread per output element, terms in a
// fixed order.
__device__ __forceinline__ uint64_t deposit(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  for (uint64_t bit = 1; mask; bit <<= 1) {
    const uint64_t low = mask & (0 - mask);
    if (v & bit) r |= low;
    mask &= mask - 1;
  }
  return r;
}

__global__ __launch_bounds__(kBlock) void ptrace_kernel(double2* __restrict__ out,
                                                       const double2* __restrict__ rho, int n,
                                                       uint32_t keep, int m) {
  const uint64_t dimk = (uint64_t)1 << m, dim = (uint64_t)1 << n, len = dimk * dimk;
  const uint64_t tr = (dim - 1) & ~(uint64_t)keep;
  const uint64_t nt = (uint64_t)1 << (n - m);
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < len;
       t += (uint64_t)gridDim.x * kBlock) {
    const uint64_t r0 = deposit(t >> m, keep), c0 = deposit(t & (dimk - 1), keep);
    double re = 0.0, im = 0.0;
    uint64_t d = 0;
    for (uint64_t u = 0; u < nt; ++u) {
      const double2 x = rho[(r0 | d) * dim + (c0 | d)];
      re += x.x;
      im += x.y;
      d = ((d | ~tr) + 1) & tr;
    }
    out[t] = double2{re, im};
  }
}

// A density of 0 qubits is the 1 x 1 matrix of a scalar state (the
// reference's empty program evolves one); a partial trace needs a qubit.
int check_n(int n, const char* what, int lo = 0) {
  if (n < lo || n > QFX_DENSITY_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, what);
  return QFX_OK;
}

}  // namespace

extern "C" {

int qfx_density_outer(void* stream, void* rho, const void* psi, int n) {
  if (!rho || !psi) return qfx::set_error(QFX_EINVAL, "qfx_density_outer: null pointer");
  if (int rc = check_n(n, "qfx_density_outer: n out of range")) return rc;
  hipLaunchKernelGGL(outer_kernel, dim3(grid_for((uint64_t)1 << (2 * n))), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double2*>(rho),
                     static_cast<const double2*>(psi), n);
  return qfx::check_launch("outer_kernel");
}

int qfx_density_identity(void* stream, void* rho, int n, double scale) {
  if (!rho) return qfx::set_error(QFX_EINVAL, "qfx_density_identity: null pointer");
  if (int rc = check_n(n, "qfx_density_identity: n out of range")) return rc;
  hipLaunchKernelGGL(identity_kernel, dim3(grid_for((uint64_t)1 << (2 * n))), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double2*>(rho), n, scale);
  return qfx::check_launch("identity_kernel");
}

int qfx_density_diag(void* stream, void* out, const void* rho, int n) {
  if (!out || !rho) return qfx::set_error(QFX_EINVAL, "qfx_density_diag: null pointer");
  if (int rc = check_n(n, "qfx_density_diag: n out of range")) return rc;
  hipLaunchKernelGGL(diag_kernel, dim3(grid_for((uint64_t)1 << n)), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double*>(out),
                     static_cast<const double2*>(rho), n);
  return qfx::check_launch("diag_kernel");
}

int qfx_density_trace(void* stream, void* out, const void* rho, int n) {
  if (!out || !rho) return qfx::set_error(QFX_EINVAL, "qfx_density_trace: null pointer");
  if (int rc = check_n(n, "qfx_density_trace: n out of range")) return rc;
  hipLaunchKernelGGL(trace_kernel, dim3(1), dim3(kTraceBlock), 0, qfx::as_stream(stream),
                     static_cast<double2*>(out), static_cast<const double2*>(rho), n);
  return qfx::check_launch("trace_kernel");
}

int qfx_density_axpy(void* stream, void* y, const void* x, double w, int64_t len) {
  if (!y || !x) return qfx::set_error(QFX_EINVAL, "qfx_density_axpy: null pointer");
  if (len < 0) return qfx::set_error(QFX_ERANGE, "qfx_density_axpy: negative length");
  if (len == 0) return QFX_OK;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for((uint64_t)len)), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double2*>(y),
                     static_cast<const double2*>(x), w, (uint64_t)len);
  return qfx::check_launch("axpy_kernel");
}

int qfx_density_partial_trace(void* stream, void* out, const void* rho, int n,
                              uint32_t keep_mask) {
  if (!out || !rho) return qfx::set_error(QFX_EINVAL, "qfx_density_partial_trace: null pointer");
  if (int rc = check_n(n, "qfx_density_partial_trace: n out of range", 1)) return rc;
  if (keep_mask == 0 || (keep_mask >> n) != 0)
    return qfx::set_error(QFX_EINVAL, "qfx_density_partial_trace: bad keep mask");
  const int m = __builtin_popcount(keep_mask);
  hipLaunchKernelGGL(ptrace_kernel, dim3(grid_for((uint64_t)1 << (2 * m))), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double2*>(out),
                     static_cast<const double2*>(rho), n, keep_mask, m);
  return qfx::check_launch("ptrace_kernel");
}

}  // extern "C"

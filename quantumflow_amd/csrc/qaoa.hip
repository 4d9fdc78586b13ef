// QAOA MaxCut kernels: the cut table (the cost diagonal).
//
// qfx_cut_table replaces graph_cuts (quantumflow/qaoa.py:68-81): for every
// basis index z, C(z) = sum over edges, in edge order, of w_e [z_u != z_v],
// node u <-> axis u <-> index bit n-1-u.  Each entry accumulates the edges in
// the reference loop's order, so the table is bitwise identical to it.
#include "qfx_common.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxEdges = 4096;

__global__ __launch_bounds__(kBlock) void cut_table_kernel(double* __restrict__ out, int n,
                                                          const int* __restrict__ edges,
                                                          const double* __restrict__ w, int m) {
  __shared__ int s_u[kMaxEdges];
  __shared__ int s_v[kMaxEdges];
  __shared__ double s_w[kMaxEdges];
  for (int e = threadIdx.x; e < m; e += kBlock) {
    s_u[e] = n - 1 - edges[2 * e];
    s_v[e] = n - 1 - edges[2 * e + 1];
    s_w[e] = w ? w[e] : 1.0;
  }
  __syncthreads();
  const uint64_t len = (uint64_t)1 << n;
  const uint64_t z = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (z >= len) return;
  double c = 0.0;
  for (int e = 0; e < m; ++e) {
    if (((z >> s_u[e]) ^ (z >> s_v[e])) & 1) c += s_w[e];
  }
  out[z] = c;
}

// Unit-weight cut counts as uint16 (sum over edges in edge order).
__global__ __launch_bounds__(kBlock) void cut_count_kernel(uint16_t* __restrict__ out, int n,
                                                          const int* __restrict__ edges, int m) {
  __shared__ int s_u[kMaxEdges];
  __shared__ int s_v[kMaxEdges];
  for (int e = threadIdx.x; e < m; e += kBlock) {
    s_u[e] = n - 1 - edges[2 * e];
    s_v[e] = n - 1 - edges[2 * e + 1];
  }
  __syncthreads();
  const uint64_t len = (uint64_t)1 << n;
  const uint64_t z = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (z >= len) return;
  unsigned c = 0;
  for (int e = 0; e < m; ++e) c += (unsigned)(((z >> s_u[e]) ^ (z >> s_v[e])) & 1);
  out[z] = (uint16_t)c;
}

}  // namespace

extern "C" {

int qfx_cut_counts(void* stream, void* out, int n, const void* edges, int m) {
  if (!out || (!edges && m > 0)) return qfx::set_error(QFX_EINVAL, "qfx_cut_counts: null pointer");
  if (n < 1 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, "qfx_cut_counts: n out of range");
  if (m < 0 || m > kMaxEdges || m > 65535)
    return qfx::set_error(QFX_ERANGE, "qfx_cut_counts: too many edges");
  const uint64_t len = (uint64_t)1 << n;
  hipLaunchKernelGGL(cut_count_kernel, dim3((unsigned)((len + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, qfx::as_stream(stream), static_cast<uint16_t*>(out), n,
                     static_cast<const int*>(edges), m);
  return qfx::check_launch("qfx_cut_counts");
}

int qfx_cut_table(void* stream, void* out, int n, const void* edges, const void* weights,
                  int m) {
  if (!out || (!edges && m > 0)) return qfx::set_error(QFX_EINVAL, "qfx_cut_table: null pointer");
  if (n < 1 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, "qfx_cut_table: n out of range");
  if (m < 0 || m > kMaxEdges) return qfx::set_error(QFX_ERANGE, "qfx_cut_table: too many edges");
  const uint64_t len = (uint64_t)1 << n;
  hipLaunchKernelGGL(cut_table_kernel, dim3((unsigned)((len + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, qfx::as_stream(stream), static_cast<double*>(out), n,
                     static_cast<const int*>(edges), static_cast<const double*>(weights), m);
  return qfx::check_launch("qfx_cut_table");
}

}  // extern "C"

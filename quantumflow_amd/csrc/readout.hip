// Readout kernels: inner product, |psi|, |psi|^2, diagonal expectation,
// all-qubit <Z>, squared norm.
//
// Replace numpybk.inner (quantumflow/backend/numpybk.py:125-128),
// State.probabilities (quantumflow/states.py:113-119), State.expectation
// (quantumflow/states.py:131-147) and QubitVector.norm
// (quantumflow/qubits.py:180-182).  Reductions run in two stages with a fixed
// work assignment (block partials, then one block in a fixed order), so a
// result is bitwise reproducible for a given length.
#include <mutex>
#include <string>

#include "qfx_common.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxRedBlocks = 1024;
constexpr int kZChunkBits = 16;  // elements per <Z> block = 2^16 (or the whole state)

// Sum `v` over the block; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /*NV*4*/) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = wave_sum(v[j]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) lds[j * 4 + wave] = v[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = ((lds[j * 4] + lds[j * 4 + 1]) + lds[j * 4 + 2]) + lds[j * 4 + 3];
  }
}

int red_blocks(uint64_t len) {
  uint64_t b = (len + kBlock * 4 - 1) / (kBlock * 4);
  if (b > (uint64_t)kMaxRedBlocks) b = kMaxRedBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

// mode 0: sum conj(a) b (complex); 1: sum |a|^2 diag; 2: sum |a|^2
template <int MODE>
__global__ __launch_bounds__(kBlock) void reduce_stage1(const double2* __restrict__ a,
                                                       const void* __restrict__ b, uint64_t len,
                                                       double2* __restrict__ partial) {
  __shared__ double lds[2 * 4];
  double acc[2] = {0.0, 0.0};
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i + 3 * stride < len; i += 4 * stride) {
    double2 x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = a[i + u * stride];
    if (MODE == 0) {
      double2 y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) y[u] = static_cast<const double2*>(b)[i + u * stride];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        acc[0] = fma(x[u].x, y[u].x, acc[0]);
        acc[0] = fma(x[u].y, y[u].y, acc[0]);
        acc[1] = fma(x[u].x, y[u].y, acc[1]);
        acc[1] = fma(-x[u].y, y[u].x, acc[1]);
      }
    } else if (MODE == 1) {
      double d[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) d[u] = static_cast<const double*>(b)[i + u * stride];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[0] = fma(fma(x[u].x, x[u].x, x[u].y * x[u].y), d[u], acc[0]);
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[0] = fma(x[u].x, x[u].x, fma(x[u].y, x[u].y, acc[0]));
    }
  }
  for (; i < len; i += stride) {
    const double2 x = a[i];
    if (MODE == 0) {
      const double2 y = static_cast<const double2*>(b)[i];
      acc[0] = fma(x.x, y.x, acc[0]);
      acc[0] = fma(x.y, y.y, acc[0]);
      acc[1] = fma(x.x, y.y, acc[1]);
      acc[1] = fma(-x.y, y.x, acc[1]);
    } else if (MODE == 1) {
      const double d = static_cast<const double*>(b)[i];
      acc[0] = fma(fma(x.x, x.x, x.y * x.y), d, acc[0]);
    } else {
      acc[0] = fma(x.x, x.x, fma(x.y, x.y, acc[0]));
    }
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(acc[0], acc[1]);
}

// Sums `nparts` partials in a fixed order; writes complex (MODE 0) or real.
template <int MODE>
__global__ __launch_bounds__(kBlock) void reduce_stage2(const double2* __restrict__ partial,
                                                       int nparts, void* __restrict__ result) {
  __shared__ double lds[2 * 4];
  double acc[2] = {0.0, 0.0};
  for (int j = threadIdx.x; j < nparts; j += kBlock) {
    acc[0] += partial[j].x;
    acc[1] += partial[j].y;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) {
    if (MODE == 0)
      static_cast<double2*>(result)[0] = make_double2(acc[0], acc[1]);
    else
      static_cast<double*>(result)[0] = acc[0];
  }
}

// partial[blk] = {sum |psi_i|^2 with bit b of i clear, ... with it set}
__global__ __launch_bounds__(kBlock) void bit_prob_stage1(const double2* __restrict__ a,
                                                         uint64_t len, int b,
                                                         double2* __restrict__ partial) {
  __shared__ double lds[2 * 4];
  double acc[2] = {0.0, 0.0};
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += stride) {
    const double2 x = a[i];
    const double p = fma(x.x, x.x, x.y * x.y);
    if ((i >> b) & 1) acc[1] += p;
    else acc[0] += p;
  }
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(acc[0], acc[1]);
}

__global__ __launch_bounds__(kBlock) void abs_kernel(double* __restrict__ out,
                                                    const double2* __restrict__ in, uint64_t len,
                                                    int squared) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= len) return;
  const double2 z = in[i];
  out[i] = squared ? fma(z.x, z.x, z.y * z.y) : hypot(z.x, z.y);
}

// <Z> stage 1: block `blk` sums chunk [blk << c, (blk+1) << c).  Element
// i = base + tid + 256 m; bits 0..7 of i are those of tid, bits 8..c-1 those
// of m.  Writes partial[blk] = {total, S_0 .. S_{c-1}} with
// S_b = sum |psi_i|^2 bit_b(i) over the chunk.
__global__ __launch_bounds__(kBlock) void z_stage1(const double2* __restrict__ psi, int c,
                                                  double* __restrict__ partial) {
  __shared__ double lds[(1 + 8 + 8) * 4];
  const uint64_t base = (uint64_t)blockIdx.x << c;
  const int tid = threadIdx.x;
  const int chunk = 1 << c;
  const int mcount = chunk > kBlock ? chunk / kBlock : 1;
  double tot = 0.0;
  double accm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (tid < chunk) {
    for (int m = 0; m < mcount; ++m) {
      const double2 z = psi[base + tid + (uint64_t)m * kBlock];
      const double p = fma(z.x, z.x, z.y * z.y);
      tot += p;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if ((m >> j) & 1) accm[j] += p;
    }
  }
  double v[17];
  v[0] = tot;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[1 + j] = ((tid >> j) & 1) ? tot : 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[9 + j] = accm[j];
  block_sum<17>(v, lds);
  if (tid == 0) {
    double* out = partial + (uint64_t)blockIdx.x * (1 + kZChunkBits);
    out[0] = v[0];
    for (int b = 0; b < kZChunkBits; ++b) {
      double s = 0.0;
      if (b < 8 && b < c) s = v[1 + b];
      else if (b >= 8 && b < c) s = v[9 + (b - 8)];
      out[1 + b] = s;
    }
  }
}

// <Z> stage 2: out[a] = total - 2 S_{n-1-a}; high bits (>= c) come from the
// block index.  One block; fixed summation order.
__global__ __launch_bounds__(kBlock) void z_stage2(const double* __restrict__ partial, int nblk,
                                                  int n, int c, double* __restrict__ out) {
  __shared__ double lds[2 * 4];
  __shared__ double s_tot;
  double acc[2] = {0.0, 0.0};
  for (int j = threadIdx.x; j < nblk; j += kBlock) acc[0] += partial[(uint64_t)j * (1 + kZChunkBits)];
  block_sum<2>(acc, lds);
  if (threadIdx.x == 0) s_tot = acc[0];
  __syncthreads();
  const double total = s_tot;
  for (int b = 0; b < n; ++b) {
    double s[2] = {0.0, 0.0};
    for (int j = threadIdx.x; j < nblk; j += kBlock) {
      const double* pj = partial + (uint64_t)j * (1 + kZChunkBits);
      if (b < c)
        s[0] += pj[1 + b];
      else if (((uint64_t)j >> (b - c)) & 1)
        s[0] += pj[0];
    }
    block_sum<2>(s, lds);
    if (threadIdx.x == 0) out[n - 1 - b] = total - 2.0 * s[0];
    __syncthreads();
  }
}

template <int MODE>
int run_reduce(hipStream_t s, void* result, const void* a, const void* b, int64_t len,
               void* ws, const char* fn) {
  if (!result || !a || !ws || (MODE != 2 && !b))
    return qfx::set_error(QFX_EINVAL, std::string(fn) + ": null pointer");
  if (len < 1) return qfx::set_error(QFX_ERANGE, std::string(fn) + ": empty input");
  const int nb = red_blocks((uint64_t)len);
  auto* part = static_cast<double2*>(ws);
  hipLaunchKernelGGL(reduce_stage1<MODE>, dim3(nb), dim3(kBlock), 0, s,
                     static_cast<const double2*>(a), b, (uint64_t)len, part);
  int rc = qfx::check_launch(fn);
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_stage2<MODE>, dim3(1), dim3(kBlock), 0, s, part, nb, result);
  return qfx::check_launch(fn);
}

}  // namespace

// ---- expectation without a caller workspace --------------------------------
// The block partials of qfx_expect_diag_noalloc live in this file-scope
// device array (16 KiB: one double2 per stage-1 block, at most
// kMaxRedBlocks), which every GPU holds once for the life of the process.
// Host code never passes the array's name to a kernel: in host code it is a
// stand-in, not a device address.  Its address on each device comes from
// hipGetSymbolAddress, taken once per device with that device current.
// Every slot stage 2 reads is written by stage 1 of the same call, so the
// array needs no clearing; calls that overlap on one GPU share it and race.
static __device__ __align__(256) double2 scratch[kMaxRedBlocks];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

int scratch_on_current_device(void** out, const char* fn) {
  int dev = 0;
  int rc = qfx::check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxScratchDevices)
    return qfx::set_error(QFX_ERANGE, std::string(fn) + ": device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    rc = qfx::check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    if (rc) return rc;
    if (!p) return qfx::set_error(QFX_EHIP, std::string(fn) + ": no address for scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return QFX_OK;
}

}  // namespace

extern "C" {

int qfx_expect_diag_noalloc(void* stream, void* result, const void* psi, const void* diag,
                            int64_t len) {
  if (len > ((int64_t)1 << QFX_MAX_QUBITS))
    return qfx::set_error(QFX_ERANGE,
                          "qfx_expect_diag_noalloc: len exceeds the maximum of 2^34 elements");
  if (!result || !psi || !diag)
    return qfx::set_error(QFX_EINVAL, "qfx_expect_diag_noalloc: null pointer");
  if (len < 1) return qfx::set_error(QFX_ERANGE, "qfx_expect_diag_noalloc: empty input");
  static_assert(sizeof(scratch) == (size_t)kMaxRedBlocks * sizeof(double2),
                "scratch holds one partial per stage-1 block");
  void* ws = nullptr;
  int rc = scratch_on_current_device(&ws, "qfx_expect_diag_noalloc");
  if (rc) return rc;
  return run_reduce<1>(qfx::as_stream(stream), result, psi, diag, len, ws,
                       "qfx_expect_diag_noalloc");
}

int64_t qfx_reduce_workspace_bytes(int64_t len) {
  int64_t z = 16;
  if (len > 1) {
    int n = 0;
    while (((int64_t)1 << n) < len) ++n;
    const int c = n < kZChunkBits ? n : kZChunkBits;
    z = ((int64_t)1 << (n - c)) * (1 + kZChunkBits) * (int64_t)sizeof(double);
  }
  const int64_t r = (int64_t)kMaxRedBlocks * (int64_t)sizeof(double2);
  return z > r ? z : r;
}

int qfx_inner(void* stream, void* result, const void* a, const void* b, int64_t len,
              void* workspace) {
  return run_reduce<0>(qfx::as_stream(stream), result, a, b, len, workspace, "qfx_inner");
}

int qfx_expect_diag(void* stream, void* result, const void* psi, const void* diag, int64_t len,
                    void* workspace) {
  return run_reduce<1>(qfx::as_stream(stream), result, psi, diag, len, workspace,
                       "qfx_expect_diag");
}

int qfx_norm2(void* stream, void* result, const void* psi, int64_t len, void* workspace) {
  return run_reduce<2>(qfx::as_stream(stream), result, psi, nullptr, len, workspace, "qfx_norm2");
}

int qfx_qubit_probs(void* stream, void* out, const void* psi, int n, int axis, void* workspace) {
  if (!out || !psi || !workspace) return qfx::set_error(QFX_EINVAL, "qfx_qubit_probs: null pointer");
  if (n < 1 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, "qfx_qubit_probs: n out of range");
  if (axis < 0 || axis >= n) return qfx::set_error(QFX_EINVAL, "qfx_qubit_probs: axis out of range");
  hipStream_t s = qfx::as_stream(stream);
  const uint64_t len = (uint64_t)1 << n;
  const int nb = red_blocks(len);
  auto* part = static_cast<double2*>(workspace);
  hipLaunchKernelGGL(bit_prob_stage1, dim3(nb), dim3(kBlock), 0, s,
                     static_cast<const double2*>(psi), len, n - 1 - axis, part);
  int rc = qfx::check_launch("qfx_qubit_probs");
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_stage2<0>, dim3(1), dim3(kBlock), 0, s, part, nb, out);
  return qfx::check_launch("qfx_qubit_probs");
}

int qfx_abs(void* stream, void* out, const void* in, int64_t len) {
  if (!out || !in) return qfx::set_error(QFX_EINVAL, "qfx_abs: null pointer");
  if (len < 1) return QFX_OK;
  hipLaunchKernelGGL(abs_kernel, dim3((unsigned)((len + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double*>(out),
                     static_cast<const double2*>(in), (uint64_t)len, 0);
  return qfx::check_launch("qfx_abs");
}

int qfx_abs2(void* stream, void* out, const void* in, int64_t len) {
  if (!out || !in) return qfx::set_error(QFX_EINVAL, "qfx_abs2: null pointer");
  if (len < 1) return QFX_OK;
  hipLaunchKernelGGL(abs_kernel, dim3((unsigned)((len + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     qfx::as_stream(stream), static_cast<double*>(out),
                     static_cast<const double2*>(in), (uint64_t)len, 1);
  return qfx::check_launch("qfx_abs2");
}

int qfx_expect_z(void* stream, void* out, const void* psi, int n, void* workspace) {
  if (!out || !psi || !workspace) return qfx::set_error(QFX_EINVAL, "qfx_expect_z: null pointer");
  if (n < 1 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, "qfx_expect_z: n out of range");
  hipStream_t s = qfx::as_stream(stream);
  const int c = n < kZChunkBits ? n : kZChunkBits;
  const uint64_t nblk = (uint64_t)1 << (n - c);
  auto* part = static_cast<double*>(workspace);
  hipLaunchKernelGGL(z_stage1, dim3((unsigned)nblk), dim3(kBlock), 0, s,
                     static_cast<const double2*>(psi), c, part);
  int rc = qfx::check_launch("qfx_expect_z");
  if (rc) return rc;
  hipLaunchKernelGGL(z_stage2, dim3(1), dim3(kBlock), 0, s, part, (int)nblk, n, c,
                     static_cast<double*>(out));
  return qfx::check_launch("qfx_expect_z");
}

}  // extern "C"

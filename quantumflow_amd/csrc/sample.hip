// Sampling the computational basis from |psi|^2 on the device.
//
// Replaces the host-side numpy draws of State.sample / State.measure
// (quantumflow/states.py:121-129 and :149-160), which copy all 2^n probabilities
// to the host: here the probabilities never leave HBM.
//
//   1. chunk sums: S_c = sum of |psi_i|^2 over chunk c (2^cb amplitudes,
//      cb = min(10, n)), in a fixed order;
//   2. exclusive prefix P_c over the chunk sums, P_nchunks = total, in a
//      fixed order (one workgroup, blocked scan);
//   3. per sample t: u = U(seed, t) * total with U a SplitMix64 stream
//      (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
//      OOPSLA 2014) mapped to [0, 1); binary search for the chunk with
//      P_c <= u < P_c+1, then one wave scans that chunk (16 amplitudes per
//      lane, wave-wide prefix) for the first index whose running sum exceeds
//      u - P_c.  Indices with probability 0 are never drawn.
// Results depend only on (psi, seed, trials): bitwise reproducible.
#include "qfx_common.h"

namespace {

constexpr int kMaxChunkBits = 10;
constexpr int kScanBlock = 1024;

__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t t) {
  uint64_t z = seed + (t + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

// one workgroup of 256 per chunk; chunk of 2^cb <= 1024 amplitudes
__global__ __launch_bounds__(256) void chunk_sum_kernel(double* __restrict__ sums,
                                                        const double2* __restrict__ psi, int cb) {
  const uint64_t c = blockIdx.x;
  const int len = 1 << cb;
  const double2* p = psi + (c << cb);
  double acc = 0.0;
  for (int i = threadIdx.x; i < len; i += 256) {
    const double2 v = p[i];
    acc = fma(v.x, v.x, fma(v.y, v.y, acc));
  }
  acc = wave_sum(acc);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sums[c] = (red[0] + red[1]) + (red[2] + red[3]);
}

// pre[0] = 0, pre[c + 1] = pre[c] + sums[c]; one workgroup, segments of 1024
__global__ __launch_bounds__(kScanBlock) void prefix_kernel(double* __restrict__ pre,
                                                            const double* __restrict__ sums,
                                                            uint64_t nchunks) {
  __shared__ double buf[kScanBlock];
  __shared__ double carry;
  if (threadIdx.x == 0) {
    carry = 0.0;
    pre[0] = 0.0;
  }
  __syncthreads();
  for (uint64_t s0 = 0; s0 < nchunks; s0 += kScanBlock) {
    const uint64_t i = s0 + threadIdx.x;
    buf[threadIdx.x] = i < nchunks ? sums[i] : 0.0;
    __syncthreads();
    // Hillis-Steele inclusive scan
    for (int d = 1; d < kScanBlock; d <<= 1) {
      const double add = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0.0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < nchunks) pre[i + 1] = carry + buf[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) carry += buf[kScanBlock - 1];
    __syncthreads();
  }
}

// one wave per sample, 4 samples per workgroup
__global__ __launch_bounds__(256) void sample_kernel(int64_t* __restrict__ out,
                                                     const double2* __restrict__ psi,
                                                     const double* __restrict__ pre,
                                                     uint64_t nchunks, int cb, int64_t trials,
                                                     uint64_t seed) {
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= trials) return;   // uniform per wave
  const double total = pre[nchunks];
  double u = uniform01(seed, (uint64_t)t) * total;
  if (u >= total) u = nextafter(total, 0.0);
  // largest c in [0, nchunks) with pre[c] <= u
  uint64_t lo = 0, hi = nchunks - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= u) lo = mid;
    else hi = mid - 1;
  }
  const double target = u - pre[lo];
  const int len = 1 << cb;
  const int per = len >= 64 ? len / 64 : 1;      // contiguous amplitudes per lane
  const double2* p = psi + (lo << cb);
  double pr[16];
  double mine = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int i = lane * per + k;
    double q = 0.0;
    if (k < per && i < len) {
      const double2 v = p[i];
      q = fma(v.x, v.x, v.y * v.y);
    }
    pr[k] = q;
    mine += q;
  }
  // inclusive wave prefix of the lane sums; the exclusive prefix is the
  // previous lane's inclusive one, so the lane intervals [excl, incl) tile
  // [0, incl of lane 63) with no gap or overlap
  double incl = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  double excl = __shfl_up(incl, 1, 64);
  if (lane == 0) excl = 0.0;
  // the one lane whose interval holds the target claims the first index
  // whose running sum passes it (its last nonzero index if rounding leaves
  // the running sum short of the target)
  int found = -1;
  if (mine > 0.0 && excl <= target && target < incl) {
    double run = excl;
    int last = -1;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k < per && pr[k] > 0.0) {
        run += pr[k];
        last = lane * per + k;
        if (found < 0 && run > target) found = last;
      }
    }
    if (found < 0) found = last;
  }
  unsigned long long claim = found >= 0 ? (unsigned long long)found : ~0ull;
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(claim, d, 64);
    claim = o < claim ? o : claim;
  }
  // no claim: the target lies at or above the wave's own chunk total
  // (rounding between the two summation orders); take the chunk's last
  // nonzero index
  if (claim == ~0ull) {
    long long last = -1;
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (k < per && pr[k] > 0.0) last = lane * per + k;
    for (int d = 32; d > 0; d >>= 1) {
      const long long o = __shfl_xor(last, d, 64);
      last = o > last ? o : last;
    }
    claim = (unsigned long long)(last >= 0 ? last : 0);
  }
  if (lane == 0) out[t] = (int64_t)((lo << cb) | claim);
}

__global__ __launch_bounds__(256) void histogram_kernel(unsigned long long* __restrict__ counts,
                                                        const int64_t* __restrict__ idx,
                                                        int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < m) atomicAdd(counts + idx[i], 1ull);
}

int chunk_bits(int n) { return n < kMaxChunkBits ? n : kMaxChunkBits; }

}  // namespace

extern "C" {

int64_t qfx_sample_workspace_bytes(int n) {
  if (n < 1 || n > QFX_MAX_QUBITS) return 0;
  const uint64_t nchunks = (uint64_t)1 << (n - chunk_bits(n));
  return (int64_t)(sizeof(double) * (2 * nchunks + 1));
}

int qfx_sample(void* stream, int64_t* out, const void* psi, int n, int64_t trials,
               uint64_t seed, void* ws) {
  if (n < 1 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, "qfx_sample: n out of range");
  if (trials < 0) return qfx::set_error(QFX_EINVAL, "qfx_sample: negative trials");
  if (trials == 0) return QFX_OK;
  if (!out || !psi || !ws) return qfx::set_error(QFX_EINVAL, "qfx_sample: null pointer");
  hipStream_t s = qfx::as_stream(stream);
  const int cb = chunk_bits(n);
  const uint64_t nchunks = (uint64_t)1 << (n - cb);
  double* sums = static_cast<double*>(ws);
  double* pre = sums + nchunks;
  auto* st = static_cast<const double2*>(psi);
  hipLaunchKernelGGL(chunk_sum_kernel, dim3((unsigned)nchunks), dim3(256), 0, s, sums, st, cb);
  int rc = qfx::check_launch("chunk_sum_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(prefix_kernel, dim3(1), dim3(kScanBlock), 0, s, pre, sums, nchunks);
  rc = qfx::check_launch("prefix_kernel");
  if (rc) return rc;
  const uint64_t blocks = ((uint64_t)trials + 3) / 4;
  hipLaunchKernelGGL(sample_kernel, dim3((unsigned)blocks), dim3(256), 0, s, out, st, pre, nchunks,
                     cb, trials, seed);
  return qfx::check_launch("sample_kernel");
}

int qfx_histogram(void* stream, int64_t* counts, const int64_t* idx, int64_t m, int n) {
  if (n < 1 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, "qfx_histogram: n out of range");
  if (m < 0) return qfx::set_error(QFX_EINVAL, "qfx_histogram: negative count");
  if (m == 0) return QFX_OK;
  if (!counts || !idx) return qfx::set_error(QFX_EINVAL, "qfx_histogram: null pointer");
  hipStream_t s = qfx::as_stream(stream);
  const uint64_t blocks = ((uint64_t)m + 255) / 256;
  hipLaunchKernelGGL(histogram_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<unsigned long long*>(counts), idx, m);
  return qfx::check_launch("histogram_kernel");
}

}  // extern "C"

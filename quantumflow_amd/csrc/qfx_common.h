// Internal helpers shared by the qfx HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/qfx.h"

namespace qfx {

// Records the error text for qfx_last_error() and returns `code`.
int set_error(int code, const std::string& msg);
// Checks the last HIP launch/call; returns QFX_OK or QFX_EHIP with text.
int check_hip(hipError_t err, const char* what);
int check_launch(const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Validates a target list: 1 <= k <= kmax, distinct, each < n.
int check_targets(int n, int k, const int* targets, int kmax, const char* fn);

// Host-side: sort k bit positions ascending into `sorted`.
void sort_bits(int k, const int* targets, int* sorted);

// Offset of gate-local index r: gate qubit j (MSB first) -> bit targets[j].
inline uint64_t local_offset(int k, const int* targets, unsigned r) {
  uint64_t off = 0;
  for (int j = 0; j < k; ++j)
    if ((r >> (k - 1 - j)) & 1u) off |= (uint64_t)1 << targets[j];
  return off;
}

}  // namespace qfx

// ---- device helpers -------------------------------------------------------

// Spread the bits of g around the (ascending) zero positions in `sorted`.
template <int K>
__device__ __forceinline__ uint64_t insert_zeros(uint64_t g, const int* sorted) {
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int b = sorted[j];
    const uint64_t low = g & (((uint64_t)1 << b) - 1);
    g = ((g >> b) << (b + 1)) | low;
  }
  return g;
}

__device__ __forceinline__ uint64_t insert_zeros_rt(uint64_t g, const int* sorted, int k) {
  for (int j = 0; j < k; ++j) {
    const int b = sorted[j];
    const uint64_t low = g & (((uint64_t)1 << b) - 1);
    g = ((g >> b) << (b + 1)) | low;
  }
  return g;
}

// Native 2-double vector: local arrays of it stay in registers (arrays of
// HIP's union-based double2 can be demoted to scratch memory).
typedef double v2d __attribute__((ext_vector_type(2)));

// Streaming loads/stores of amplitudes that are touched once per pass:
// nontemporal, so they do not displace reusable lines in the caches.
__device__ __forceinline__ v2d ld_stream(const v2d* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_stream(v2d* p, v2d v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ double2 ld_stream(const double2* p) {
  const v2d v = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(double2* p, double2 v) {
  __builtin_nontemporal_store(v2d{v.x, v.y}, reinterpret_cast<v2d*>(p));
}

__device__ __forceinline__ void cmac(v2d& acc, const v2d a, const v2d b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

__device__ __forceinline__ v2d cmul(const v2d a, const v2d b) {
  v2d r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// complex multiply-accumulate: acc += a * b
__device__ __forceinline__ void cmac(double2& acc, const double2 a, const double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

__device__ __forceinline__ double2 cmul(const double2 a, const double2 b) {
  double2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// Wave64 sum of a double.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Kernels of the MLP angle proposer (fp32): forward and backward of dense
// layers, ReLU, mean-squared-error loss and Adam.  No reference counterpart:
// the reference has no learned component; the proposer warm-starts the QAOA
// optimizer of examples/qaoa_maxcut.py:37-87.
//
// Dense layer y = relu(x W^T + b), x [B, in], W [out, in] (row-major).
// All products go through one LDS-tiled GEMM with strided operand access:
//   C[m][n] = sum_k A(m, k) B(k, n),  A(m, k) = A[m*sam + k*sak],
//                                     B(k, n) = B[k*sbk + n*sbn]
// forward: (x, W^T); input gradient: (dz, W); weight gradient: (dz^T, x).
// Every output element is summed by one lane in a fixed k order, so results
// are bitwise reproducible.
#include "qfx_common.h"

namespace {

constexpr int TM = 64, TN = 64, TK = 16;  // C tile per workgroup, k step
constexpr int kThreads = 256;             // 16 x 16 lanes, 4 x 4 outputs each

__global__ __launch_bounds__(kThreads) void gemm_kernel(
    float* __restrict__ C, const float* __restrict__ A, const float* __restrict__ Bm, int M,
    int N, int K, long sam, long sak, long sbk, long sbn, const float* __restrict__ bias,
    int relu) {
  __shared__ float sA[TK][TM + 4];
  __shared__ float sB[TK][TN + 4];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += TK) {
    // stage A (TM x TK) and B (TK x TN): 1024 elements each, 4 per lane
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = threadIdx.x + q * kThreads;
      const int kk = idx % TK, mm = idx / TK;
      const int m = m0 + mm, k = k0 + kk;
      sA[kk][mm] = (m < M && k < K) ? A[(long)m * sam + (long)k * sak] : 0.0f;
      const int nn = idx % TN, kb = idx / TN;
      const int n = n0 + nn, k2 = k0 + kb;
      sB[kb][nn] = (n < N && k2 < K) ? Bm[(long)k2 * sbk + (long)n * sbn] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sA[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sB[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + ty + 16 * i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tx + 16 * j;
      if (n >= N) continue;
      float v = acc[i][j];
      if (bias) v += bias[n];
      if (relu) v = v > 0.0f ? v : 0.0f;
      C[(long)m * N + n] = v;
    }
  }
}

// dz = dy * (y > 0)
__global__ void relu_bwd_kernel(float* __restrict__ dz, const float* __restrict__ dy,
                                const float* __restrict__ y, long len) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) dz[i] = y[i] > 0.0f ? dy[i] : 0.0f;
}

// out[n] = sum_b x[b][n], one lane per column, rows in order
__global__ void colsum_kernel(float* __restrict__ out, const float* __restrict__ x, int B, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) s += x[(long)b * N + n];
  out[n] = s;
}

// dy = 2 (y - t) / (B N); per-block partial sums of (y - t)^2 / (B N)
__global__ void mse_kernel(float* __restrict__ dy, float* __restrict__ partial,
                           const float* __restrict__ y, const float* __restrict__ t, long len,
                           float scale) {
  __shared__ float red[4];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.0f;
  if (i < len) {
    const float d = y[i] - t[i];
    dy[i] = 2.0f * d * scale;
    sq = d * d * scale;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ void sum_kernel(float* __restrict__ out, const float* __restrict__ x, int len) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.0f;
    for (int i = 0; i < len; ++i) s += x[i];
    out[0] = s;
  }
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, long len, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const float gi = g[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

int grid1(long len, int bs) { return (int)((len + bs - 1) / bs); }

}  // namespace

extern "C" {

int qfx_mlp_gemm(void* stream, void* C, const void* A, const void* B, int M, int N, int K,
                 int64_t sam, int64_t sak, int64_t sbk, int64_t sbn, const void* bias, int relu) {
  if (!C || !A || !B) return qfx::set_error(QFX_EINVAL, "qfx_mlp_gemm: null pointer");
  if (M < 1 || N < 1 || K < 1) return qfx::set_error(QFX_ERANGE, "qfx_mlp_gemm: empty product");
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(kThreads), 0, qfx::as_stream(stream),
                     static_cast<float*>(C), static_cast<const float*>(A),
                     static_cast<const float*>(B), M, N, K, (long)sam, (long)sak, (long)sbk,
                     (long)sbn, static_cast<const float*>(bias), relu);
  return qfx::check_launch("qfx_mlp_gemm");
}

int qfx_mlp_relu_bwd(void* stream, void* dz, const void* dy, const void* y, int64_t len) {
  if (!dz || !dy || !y) return qfx::set_error(QFX_EINVAL, "qfx_mlp_relu_bwd: null pointer");
  if (len < 1) return QFX_OK;
  hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid1(len, 256)), dim3(256), 0, qfx::as_stream(stream),
                     static_cast<float*>(dz), static_cast<const float*>(dy),
                     static_cast<const float*>(y), (long)len);
  return qfx::check_launch("qfx_mlp_relu_bwd");
}

int qfx_mlp_colsum(void* stream, void* out, const void* x, int B, int N) {
  if (!out || !x) return qfx::set_error(QFX_EINVAL, "qfx_mlp_colsum: null pointer");
  if (B < 1 || N < 1) return qfx::set_error(QFX_ERANGE, "qfx_mlp_colsum: empty input");
  hipLaunchKernelGGL(colsum_kernel, dim3(grid1(N, 256)), dim3(256), 0, qfx::as_stream(stream),
                     static_cast<float*>(out), static_cast<const float*>(x), B, N);
  return qfx::check_launch("qfx_mlp_colsum");
}

int64_t qfx_mlp_mse_workspace(int64_t len) { return ((len + 255) / 256) * (int64_t)sizeof(float); }

int qfx_mlp_mse(void* stream, void* dy, void* loss, const void* y, const void* t, int64_t len,
                void* workspace) {
  if (!dy || !loss || !y || !t || !workspace)
    return qfx::set_error(QFX_EINVAL, "qfx_mlp_mse: null pointer");
  if (len < 1) return qfx::set_error(QFX_ERANGE, "qfx_mlp_mse: empty input");
  hipStream_t s = qfx::as_stream(stream);
  const int nb = grid1(len, 256);
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(256), 0, s, static_cast<float*>(dy),
                     static_cast<float*>(workspace), static_cast<const float*>(y),
                     static_cast<const float*>(t), (long)len, 1.0f / (float)len);
  int rc = qfx::check_launch("qfx_mlp_mse");
  if (rc) return rc;
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, s, static_cast<float*>(loss),
                     static_cast<const float*>(workspace), nb);
  return qfx::check_launch("qfx_mlp_mse");
}

int qfx_mlp_adam(void* stream, void* params, const void* grads, void* m, void* v, int64_t len,
                 float lr, float beta1, float beta2, float eps, int step, float weight_decay) {
  if (!params || !grads || !m || !v) return qfx::set_error(QFX_EINVAL, "qfx_mlp_adam: null pointer");
  if (step < 1) return qfx::set_error(QFX_EINVAL, "qfx_mlp_adam: step counts from 1");
  if (len < 1) return QFX_OK;
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(grid1(len, 256)), dim3(256), 0, qfx::as_stream(stream),
                     static_cast<float*>(params), static_cast<const float*>(grads),
                     static_cast<float*>(m), static_cast<float*>(v), (long)len, lr, beta1, beta2,
                     eps, c1, c2, weight_decay);
  return qfx::check_launch("qfx_mlp_adam");
}

}  // extern "C"

// LDS-tiled multi-gate passes: the fused circuit executor.
//
// One pass streams the whole state through LDS once.  Workgroup b owns the
// "tile" of 2^T amplitudes whose index bits outside the pass's tile bit set
// S = {s_0 < ... < s_{T-1}} spell b; it loads the tile into LDS (coalesced:
// S always holds the lowest index bits), applies every op of the pass to it
// in LDS -- each op a dense, diagonal or monomial gate on up to 5 tile
// qubits -- and writes it back in place.  Gates that act only inside S are
// thereby applied for one HBM read + write of the state instead of one each.
//
// Replaces the sequential per-gate loop of Circuit.run
// (quantumflow/circuits.py:87-99) for runs of gates the planner
// (quantumflow_amd/fusion.py) groups into passes.
#include <cstring>

#include "qfx_common.h"

namespace {

enum OpKind : int { OP_DENSE = 0, OP_SPARSE = 1 };

struct TilePos {
  int s[QFX_MAX_QUBITS];  // tile bit positions, ascending
};

__device__ __forceinline__ unsigned ins0(unsigned g, int b) {
  return ((g >> b) << (b + 1)) | (g & ((1u << b) - 1u));
}

// Local (tile) offset of gate-local index r: gate qubit j (MSB first) is
// tile bit t[j].
template <int K>
__device__ __forceinline__ unsigned loc_off(const int* t, unsigned r) {
  unsigned o = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) o |= ((r >> (K - 1 - j)) & 1u) << t[j];
  return o;
}

template <int K, int BS>
__device__ void op_dense(double2* lds, int T, const qfx_tile_op& op,
                         const double2* __restrict__ coef) {
  constexpr int D = 1 << K;
  const double2* m = coef + op.coef;
  const int ngroups = 1 << (T - K);
  for (int g = threadIdx.x; g < ngroups; g += BS) {
    unsigned lb = (unsigned)g;
#pragma unroll
    for (int j = 0; j < K; ++j) lb = ins0(lb, op.sorted[j]);
    double2 v[D];
#pragma unroll
    for (int c = 0; c < D; ++c) v[c] = lds[lb + loc_off<K>(op.t, c)];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < D; ++c) cmac(acc, m[r * D + c], v[c]);
      lds[lb + loc_off<K>(op.t, r)] = acc;
    }
  }
}

template <int K, int BS>
__device__ void op_sparse(double2* lds, int T, const qfx_tile_op& op,
                          const double2* __restrict__ coef, const int* __restrict__ srcs) {
  constexpr int D = 1 << K;
  const double2* ph = coef + op.coef;
  const int* src = srcs + op.src;
  const int ngroups = 1 << (T - K);
  for (int g = threadIdx.x; g < ngroups; g += BS) {
    unsigned lb = (unsigned)g;
#pragma unroll
    for (int j = 0; j < K; ++j) lb = ins0(lb, op.sorted[j]);
    double2 v[D];
#pragma unroll
    for (int r = 0; r < D; ++r)
      if (op.active & (1u << r)) v[r] = lds[lb + loc_off<K>(op.t, (unsigned)src[r])];
#pragma unroll
    for (int r = 0; r < D; ++r)
      if (op.active & (1u << r))
        lds[lb + loc_off<K>(op.t, r)] = (op.phase & (1u << r)) ? cmul(ph[r], v[r]) : v[r];
  }
}

template <int BS>
__device__ __forceinline__ void apply_op(double2* lds, int T, const qfx_tile_op& op,
                                         const double2* __restrict__ coef,
                                         const int* __restrict__ srcs) {
  if (op.kind == OP_DENSE) {
    switch (op.k) {
      case 1: op_dense<1, BS>(lds, T, op, coef); break;
      case 2: op_dense<2, BS>(lds, T, op, coef); break;
      case 3: op_dense<3, BS>(lds, T, op, coef); break;
      case 4: op_dense<4, BS>(lds, T, op, coef); break;
      case 5: op_dense<5, BS>(lds, T, op, coef); break;
    }
  } else {
    switch (op.k) {
      case 1: op_sparse<1, BS>(lds, T, op, coef, srcs); break;
      case 2: op_sparse<2, BS>(lds, T, op, coef, srcs); break;
      case 3: op_sparse<3, BS>(lds, T, op, coef, srcs); break;
      case 4: op_sparse<4, BS>(lds, T, op, coef, srcs); break;
      case 5: op_sparse<5, BS>(lds, T, op, coef, srcs); break;
    }
  }
}

template <int BS, int LB>
__global__ __launch_bounds__(BS) void tile_pass_kernel(double2* state, int T, TilePos pos,
                                                      const qfx_tile_op* __restrict__ ops,
                                                      int nops, const double2* __restrict__ coef,
                                                      const int* __restrict__ srcs) {
  extern __shared__ double2 lds[];
  const int tid = threadIdx.x;
  // tile base: block index bits spread over the non-tile positions
  uint64_t base = blockIdx.x;
  for (int j = 0; j < T; ++j) {
    const int b = pos.s[j];
    base = ((base >> b) << (b + 1)) | (base & (((uint64_t)1 << b) - 1));
  }
  uint64_t toff = 0;
#pragma unroll
  for (int j = 0; j < LB; ++j) toff |= (uint64_t)((tid >> j) & 1) << pos.s[j];
  const int per = 1 << (T - LB);
  const double2* src = state + base + toff;
  for (int kk = 0; kk < per; ++kk) {
    uint64_t koff = 0;
    for (int j = 0; j < T - LB; ++j) koff |= (uint64_t)((kk >> j) & 1) << pos.s[LB + j];
    lds[tid + kk * BS] = src[koff];
  }
  __syncthreads();
  for (int i = 0; i < nops; ++i) {
    apply_op<BS>(lds, T, ops[i], coef, srcs);
    __syncthreads();
  }
  double2* dst = state + base + toff;
  for (int kk = 0; kk < per; ++kk) {
    uint64_t koff = 0;
    for (int j = 0; j < T - LB; ++j) koff |= (uint64_t)((kk >> j) & 1) << pos.s[LB + j];
    dst[koff] = lds[tid + kk * BS];
  }
}

template <int BS, int LB>
int launch_pass(hipStream_t s, double2* state, int n, int T, const TilePos& pos,
                const qfx_tile_op* ops, int nops, const double2* coef, const int* srcs) {
  const size_t lds = sizeof(double2) << T;
  static bool attr_set = false;
  if (!attr_set) {
    int rc = qfx::check_hip(hipFuncSetAttribute((const void*)tile_pass_kernel<BS, LB>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                160 * 1024),
                            "hipFuncSetAttribute");
    if (rc) return rc;
    attr_set = true;
  }
  const uint64_t tiles = (uint64_t)1 << (n - T);
  hipLaunchKernelGGL((tile_pass_kernel<BS, LB>), dim3((unsigned)tiles), dim3(BS), lds, s, state, T,
                     pos, ops, nops, coef, srcs);
  return qfx::check_launch("tile_pass_kernel");
}

int check_pass(int n, int T, const int* tile_pos, int nops, int block, const char* fn) {
  if (n < 6 || n > QFX_MAX_QUBITS) return qfx::set_error(QFX_ERANGE, std::string(fn) + ": n out of range");
  if (T < 6 || T > 13 || T > n)
    return qfx::set_error(QFX_ERANGE, std::string(fn) + ": tile width must be 6..13 and <= n");
  if (!tile_pos) return qfx::set_error(QFX_EINVAL, std::string(fn) + ": null tile positions");
  for (int j = 0; j < T; ++j) {
    if (tile_pos[j] < 0 || tile_pos[j] >= n || (j > 0 && tile_pos[j] <= tile_pos[j - 1]))
      return qfx::set_error(QFX_EINVAL,
                            std::string(fn) + ": tile positions must be ascending and < n");
  }
  if (nops < 0) return qfx::set_error(QFX_EINVAL, std::string(fn) + ": negative op count");
  if (block != 64 && block != 256 && block != 512 && block != 1024)
    return qfx::set_error(QFX_EINVAL, std::string(fn) + ": block must be 64, 256, 512 or 1024");
  const int lb = block == 64 ? 6 : (block == 256 ? 8 : (block == 512 ? 9 : 10));
  if (T < lb) return qfx::set_error(QFX_ERANGE, std::string(fn) + ": tile smaller than the block");
  return QFX_OK;
}

}  // namespace

extern "C" {

int qfx_tile_check_ops(int T, const qfx_tile_op* ops, int nops, int ncoef, const int* srcs,
                       int nsrcs) {
  for (int i = 0; i < nops; ++i) {
    const qfx_tile_op& o = ops[i];
    if ((o.kind != OP_DENSE && o.kind != OP_SPARSE) || o.k < 1 || o.k > 5 || o.k > T)
      return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: bad op kind or width");
    const int D = 1 << o.k;
    const int need = o.kind == OP_DENSE ? D * D : D;
    if (o.coef < 0 || o.coef + need > ncoef)
      return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: coefficients out of range");
    for (int j = 0; j < o.k; ++j) {
      if (o.t[j] < 0 || o.t[j] >= T)
        return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: target outside the tile");
      if (j > 0 && o.sorted[j] <= o.sorted[j - 1])
        return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: sorted targets not ascending");
      bool found = false;
      for (int i2 = 0; i2 < o.k; ++i2) found |= o.sorted[i2] == o.t[j];
      if (!found) return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: sorted != targets");
    }
    if (o.kind == OP_SPARSE) {
      if (o.src < 0 || o.src + D > nsrcs)
        return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: source rows out of range");
      unsigned seen = 0;
      for (int r = 0; r < D; ++r) {
        const int sr = srcs[o.src + r];
        if (sr < 0 || sr >= D || (seen >> sr) & 1u)
          return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: source rows not a permutation");
        seen |= 1u << sr;
      }
    }
  }
  return QFX_OK;
}

int qfx_tile_pass(void* stream, void* state, int n, int tile_bits, const int* tile_pos,
                  const void* ops, int nops, const void* coef, const void* srcs, int block) {
  int rc = check_pass(n, tile_bits, tile_pos, nops, block, "qfx_tile_pass");
  if (rc) return rc;
  if (!state || (nops > 0 && (!ops || !coef)))
    return qfx::set_error(QFX_EINVAL, "qfx_tile_pass: null pointer");
  TilePos pos;
  std::memset(&pos, 0, sizeof(pos));
  for (int j = 0; j < tile_bits; ++j) pos.s[j] = tile_pos[j];
  hipStream_t s = qfx::as_stream(stream);
  auto* st = static_cast<double2*>(state);
  auto* o = static_cast<const qfx_tile_op*>(ops);
  auto* c = static_cast<const double2*>(coef);
  auto* sr = static_cast<const int*>(srcs);
  switch (block) {
    case 64: return launch_pass<64, 6>(s, st, n, tile_bits, pos, o, nops, c, sr);
    case 256: return launch_pass<256, 8>(s, st, n, tile_bits, pos, o, nops, c, sr);
    case 512: return launch_pass<512, 9>(s, st, n, tile_bits, pos, o, nops, c, sr);
    default: return launch_pass<1024, 10>(s, st, n, tile_bits, pos, o, nops, c, sr);
  }
}

}  // extern "C"

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
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>

#include "qfx_common.h"

namespace {

enum OpKind : int { OP_DENSE = 0, OP_SPARSE = 1, OP_RPHASE = 2, OP_MFMA = 3 };

constexpr int kLoadBatch = 16;  // global loads in flight per lane
constexpr int kMaxTileBits = 13;
constexpr int kWideOpMaxBlock = 256;  // largest block running 4/5-qubit ops
constexpr int kRphaseMaxBlock = 512;  // largest block running register phases
constexpr int kMfmaMaxBlock = 512;    // largest block running matrix-core blocks
// Passes without 4-5 qubit ops or matrix-core blocks copy their coefficient
// table (at most kMaxStaged entries) into LDS behind the tile once per
// workgroup: uniform coefficient reads then come from LDS broadcasts instead
// of scalar registers, which a pass's 2x2/4x4 matrices and 16-slot phase
// vectors would otherwise exhaust.  12 KiB keeps two 2^12-amplitude tiles per
// CU (2 x 76 KiB of the 160 KiB).
constexpr int kMaxStaged = 768;

// Where a pass loads and stores.  A pass is in place: it stores each tile
// to the addresses it loaded, possibly with its tile bits permuted (tile bit
// j is stored to index bit d[j], a permutation of s[0..T-1]: the qubit
// relabelling of the SWAP gates the pass absorbed).
struct TilePos {
  int s[QFX_MAX_QUBITS];   // tile bit positions, ascending (where tile bit j loads from)
  int d[QFX_MAX_QUBITS];   // where tile bit j is stored
  int permuted;            // d differs from s somewhere
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

// ---- LDS layout -----------------------------------------------------------
// Element e of a tile lives at LDS slot sw(e): the 3-bit chunks above bit 2
// are folded into bits 0..2, so lanes whose elements differ only in higher
// bits (strided amplitude groups) still hit different LDS banks.
__host__ __device__ __forceinline__ unsigned sw(unsigned e) {
  return e ^ (((e >> 3) ^ (e >> 6) ^ (e >> 9) ^ (e >> 12)) & 7u);
}

// Offset of register index r when register bit j sits at tile bit t[j].
template <int K>
__device__ __forceinline__ unsigned loc_off_lsb(const int* t, unsigned r) {
  unsigned o = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) o |= ((r >> j) & 1u) << t[j];
  return o;
}

// groups per lane per iteration: keep ~8 LDS reads in flight
template <int K>
constexpr int op_groups() { return K == 1 ? 4 : (K == 2 ? 2 : 1); }

// Coefficient pointers of the LDS ops are deliberately not __restrict__:
// the coefficients may then not be hoisted above the op's own LDS stores,
// so a matrix is re-read per group (cheap uniform reads) instead of pinning
// up to 4^k complex values in registers for the whole op.
template <int K, int BS>
__device__ __forceinline__ void op_dense(v2d* lds, int T, const qfx_tile_op& op,
                         const v2d* coef) {
  constexpr int D = 1 << K;
  constexpr int G = op_groups<K>();
  const v2d* m = coef + op.coef;
  const int ngroups = 1 << (T - K);   // a multiple of BS * G for T >= 10
  for (int g0 = threadIdx.x; g0 < ngroups; g0 += BS * G) {
    unsigned lb[G];
    v2d v[G][D];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      unsigned l = (unsigned)(g0 + u * BS);
#pragma unroll
      for (int j = 0; j < K; ++j) l = ins0(l, op.sorted[j]);
      lb[u] = l;
      if (g0 + u * BS < ngroups) {
#pragma unroll
        for (int c = 0; c < D; ++c) v[u][c] = lds[sw(l + loc_off<K>(op.t, c))];
      }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (g0 + u * BS < ngroups) {
#pragma unroll
        for (int r = 0; r < D; ++r) {
          v2d acc = {0.0, 0.0};
#pragma unroll
          for (int c = 0; c < D; ++c) cmac(acc, m[r * D + c], v[u][c]);
          lds[sw(lb[u] + loc_off<K>(op.t, r))] = acc;
          // one matrix row in registers at a time
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  }
}

template <int K, int BS>
__device__ __forceinline__ void op_sparse(v2d* lds, int T, const qfx_tile_op& op,
                          const v2d* coef, const int* __restrict__ srcs) {
  constexpr int D = 1 << K;
  constexpr int G = op_groups<K>();
  const v2d* ph = coef + op.coef;
  const int* src = srcs + op.src;
  const int ngroups = 1 << (T - K);
  for (int g0 = threadIdx.x; g0 < ngroups; g0 += BS * G) {
    unsigned lb[G];
    v2d v[G][D];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      unsigned l = (unsigned)(g0 + u * BS);
#pragma unroll
      for (int j = 0; j < K; ++j) l = ins0(l, op.sorted[j]);
      lb[u] = l;
      if (g0 + u * BS < ngroups) {
#pragma unroll
        for (int r = 0; r < D; ++r)
          if (op.active & (1u << r)) v[u][r] = lds[sw(l + loc_off<K>(op.t, (unsigned)src[r]))];
      }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (g0 + u * BS < ngroups) {
#pragma unroll
        for (int r = 0; r < D; ++r)
          if (op.active & (1u << r))
            lds[sw(lb[u] + loc_off<K>(op.t, r))] =
                (op.phase & (1u << r)) ? cmul(ph[r], v[u][r]) : v[u][r];
      }
    }
  }
}


// ---- register phases ------------------------------------------------------
// A register phase holds, per lane, the 16 amplitudes of one group spanned by
// 4 tile bits (register bit j <-> tile bit op.t[j]) and applies a run of
// gates to them in registers, so the run costs one LDS sweep.
// Sub-ops (4 ints each in `subs`): {type, code, coef index, mask}.
//   SUB_D1  2x2 matrix on register bit `code`
//   SUB_D1R 2x2 real matrix (real parts of the 4 coefficients; RY, H)
//   SUB_D1X 2x2 matrix with real diagonal and imaginary off-diagonal
//           (RX-like: real parts of m00, m11, imaginary parts of m01, m10)
//   SUB_DIAG  16 phases (one per slot; exactly 1 where `mask` has no bit)
//   SUB_DIAGV 2^m tables of 16 phases, m = code & 3 (1..3): table v where
//           bit j of v is selector p = (code >> (2 + 7 j)) & 127: state-index
//           bit p (< 64, outside the tile) or, for p >= 64, tile bit p - 64
//           of the lane group (a non-register bit of the phase).  Diagonal
//           gates thus need none of their qubits in registers.  With
//           `mask` = 16 | d, the tables are compact: they depend only on the
//           register bits in d, table v holds 2^popcount(d) phases, and slot
//           r reads the entry indexed by r's bits under d, in order.
//   SUB_CX  X on register bit t = code & 3 for the slots whose bits under
//           the control mask (code >> 2) & 15 equal (code >> 6) & 15: a pure
//           register exchange (CNOT, CCNOT; controlled SWAP as three).
//           `mask` may add up to 3 controls that are not register bits:
//           m = mask & 3 selectors coded as for SUB_DIAGV, selector j
//           required to equal bit 23 + j of `mask`; where they do not, the
//           sub-op leaves the lane's amplitudes as they are.  Controlled-X
//           gates thus need only their target in registers.
// Any record may also carry lane flips (see kD1Flip below).
// Uncontrolled exchanges (SWAP, X, the permutation of ISWAP and Y) never
// reach the kernel: the host relabels which logical slot each register
// holds, and a leading run of SWAP/X/CNOT gates is folded into the load.
// The phase's load map (op.phase) and store map (op.active) are affine maps
// of the 4-bit logical index, matrix | offset << 16 with column j the image
// of bit j: register r loads from and stores to the slots of their images.
// Exchanges move amplitudes bit for bit; phases of monomial gates follow as
// a SUB_DIAG.  Dense 2-qubit gates run as LDS ops: a 4x4
// matrix next to 16 register amplitudes would raise the register allocation
// of the whole kernel past two waves per SIMD.
enum SubKind : int {
  SUB_D1 = 1, SUB_DIAG = 4, SUB_CX = 5, SUB_D1R = 7, SUB_D1X = 8, SUB_DIAGV = 9
};
constexpr int kRegBits = 4;
constexpr int kRegs = 1 << kRegBits;

// Lane flips.  A controlled X whose controls are all selectors (index bits
// outside the tile, lane-group tile bits) moves no data: the lane records
// it in a flip mask fm over its register bits, so that register r holds
// what register r ^ fm holds on a lane without flips, and the phase's store
// writes register r where register r ^ fm belongs (one XOR of its address
// base).  A flip rides on a later record, as a 20-bit word in bits 8.. of
// its type word, applied before the record.  Word: bit 0 set,
// bits 1-2 register bit, bits 3-9 selector code (as SUB_DIAGV's) and bit 10
// its required value, bit 11 a second selector in bits 12-18 with its value
// in bit 19.  The host folds flips into the selector tables of later
// diagonals; a 2x2 matrix on a flipped bit carries kD1Flip and 8
// coefficients (m, then X m X) and reads the second half on flipped lanes;
// a CX whose controls are flipped carries kCxFlip; a diagonal names in its
// SUB_DIAGV mask bits 8.. the flip-mask bits it reads, and its slot r takes
// the entry of slot r ^ (fm & those bits).
constexpr int kD1Flip = 4;                 // 2x2 codes
constexpr int kCxFlip = 1 << 10;           // SUB_CX code: controls compared through fm
constexpr int kRecType = 255;

// 1-qubit sub-ops walk the 8 register pairs (r, r | 1 << A); a scheduling
// fence after the first four keeps half of the results in flight at a time
// (all eight would push the kernel past 256 VGPRs, i.e. below two waves per
// SIMD).
template <int A>
constexpr int pair_lo(int p) { return ((p >> A) << (A + 1)) | (p & ((1 << A) - 1)); }

template <int A>
__device__ __forceinline__ void sub_d1(v2d (&v)[kRegs], const v2d m0, const v2d m1, const v2d m2,
                                       const v2d m3) {
#pragma unroll
  for (int p = 0; p < kRegs / 2; ++p) {
    const int r = pair_lo<A>(p);
    const v2d x = v[r], y = v[r | (1 << A)];
    v2d a = {0.0, 0.0}, b = {0.0, 0.0};
    cmac(a, m0, x);
    cmac(a, m1, y);
    cmac(b, m2, x);
    cmac(b, m3, y);
    v[r] = a;
    v[r | (1 << A)] = b;
    if (p == kRegs / 4 - 1) __builtin_amdgcn_sched_barrier(0);
  }
}

template <int A>
__device__ __forceinline__ void sub_d1r(v2d (&v)[kRegs], const double m0, const double m1,
                                        const double m2, const double m3) {
#pragma unroll
  for (int p = 0; p < kRegs / 2; ++p) {
    const int r = pair_lo<A>(p);
    const v2d x = v[r], y = v[r | (1 << A)];
    v[r] = v2d{fma(m0, x.x, m1 * y.x), fma(m0, x.y, m1 * y.y)};
    v[r | (1 << A)] = v2d{fma(m2, x.x, m3 * y.x), fma(m2, x.y, m3 * y.y)};
    if (p == kRegs / 4 - 1) __builtin_amdgcn_sched_barrier(0);
  }
}

// [[c0, i s0], [i s1, c1]]
template <int A>
__device__ __forceinline__ void sub_d1x(v2d (&v)[kRegs], const double c0, const double s0,
                                        const double s1, const double c1) {
#pragma unroll
  for (int p = 0; p < kRegs / 2; ++p) {
    const int r = pair_lo<A>(p);
    const v2d x = v[r], y = v[r | (1 << A)];
    v[r] = v2d{fma(c0, x.x, -s0 * y.y), fma(c0, x.y, s0 * y.x)};
    v[r | (1 << A)] = v2d{fma(c1, y.x, -s1 * x.y), fma(c1, y.y, s1 * x.x)};
    if (p == kRegs / 4 - 1) __builtin_amdgcn_sched_barrier(0);
  }
}

// Every slot is multiplied (slots outside `mask` hold exactly 1, which
// leaves an amplitude bit for bit unchanged): no per-slot branches.
__device__ __forceinline__ void sub_diag(v2d (&v)[kRegs], const v2d* __restrict__ ph) {
#pragma unroll
  for (int r0 = 0; r0 < kRegs; r0 += 4) {
#pragma unroll
    for (int r = r0; r < r0 + 4; ++r) v[r] = cmul(ph[r], v[r]);
    __builtin_amdgcn_sched_barrier(0);   // four phases in registers at a time
  }
}

// A diagonal given by a compact table over the register bits in `dep`:
// slot r multiplies by ph[index of r's bits under dep], the index a sum of
// per-bit strides (uniform scalars), so no register is indexed at run time.
__device__ __forceinline__ void sub_diag_compact(v2d (&v)[kRegs], const v2d* __restrict__ ph,
                                                 unsigned dep) {
  unsigned stride[kRegBits];
  unsigned k = 0;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) {
    stride[j] = ((dep >> j) & 1u) ? (1u << k) : 0u;
    k += (dep >> j) & 1u;
  }
#pragma unroll
  for (int r0 = 0; r0 < kRegs; r0 += 4) {
#pragma unroll
    for (int r = r0; r < r0 + 4; ++r) {
      unsigned idx = 0;
#pragma unroll
      for (int j = 0; j < kRegBits; ++j)
        if (r & (1 << j)) idx += stride[j];
      v[r] = cmul(ph[idx], v[r]);
    }
    __builtin_amdgcn_sched_barrier(0);   // four phases in registers at a time
  }
}

// Diagonals on a lane with flips f: slot r takes the phase of slot r ^ f.
__device__ __forceinline__ void sub_diag_flip(v2d (&v)[kRegs], const v2d* __restrict__ ph,
                                              unsigned f) {
#pragma unroll
  for (int r0 = 0; r0 < kRegs; r0 += 4) {
#pragma unroll
    for (int r = r0; r < r0 + 4; ++r) v[r] = cmul(ph[(unsigned)r ^ f], v[r]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// The compact index is XOR-linear in the slot (distinct bits per register
// bit), so slot r ^ f has index idx(r) ^ idx(f).
__device__ __forceinline__ void sub_diag_compact_flip(v2d (&v)[kRegs], const v2d* __restrict__ ph,
                                                      unsigned dep, unsigned f) {
  unsigned stride[kRegBits];
  unsigned k = 0, fi = 0;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) {
    stride[j] = ((dep >> j) & 1u) ? (1u << k) : 0u;
    k += (dep >> j) & 1u;
    fi |= ((f >> j) & 1u) ? stride[j] : 0u;
  }
#pragma unroll
  for (int r0 = 0; r0 < kRegs; r0 += 4) {
#pragma unroll
    for (int r = r0; r < r0 + 4; ++r) {
      unsigned idx = 0;
#pragma unroll
      for (int j = 0; j < kRegBits; ++j)
        if (r & (1 << j)) idx += stride[j];
      v[r] = cmul(ph[idx ^ fi], v[r]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Marks a branch that reads flips: the compiler may not turn its few
// instructions into selects that every record without the flag would run.
__device__ __forceinline__ void flagged_branch() { __asm__ volatile(""); }

// Whether the selector controls of word w (as in SUB_CX) hold.
__device__ __forceinline__ bool selectors_hold(unsigned w, uint64_t gbase, unsigned egrp) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j < (int)(w & 3u)) {
      const unsigned p = (w >> (2 + 7 * j)) & 127u;
      const unsigned b = p < 64u ? (unsigned)((gbase >> p) & 1u) : (egrp >> (p - 64u)) & 1u;
      ok = ok && b == ((w >> (23 + j)) & 1u);
    }
  }
  return ok;
}

// Coefficient offset of a 2x2 record on this lane: with selectors in its
// mask word (coded as SUB_DIAGV's code) the record holds one set of
// coefficients per selector value v, 4 each, 8 when flagged (m, then
// X m X, the second read where the bit is flipped).
__device__ __forceinline__ unsigned d1_offset(const int4 sub, unsigned fm, uint64_t gbase,
                                              unsigned egrp) {
  const int code = sub.y;
  const bool flag = code & kD1Flip;
  unsigned off = flag ? (((fm >> (code & 3)) & 1u) << 2) : 0u;
  const unsigned w = (unsigned)sub.w;
  if (w) {                                  // uniform: most records have none
    flagged_branch();
    unsigned sel = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (j < (int)(w & 3u)) {
        const unsigned p = (w >> (2 + 7 * j)) & 127u;
        const unsigned b = p < 64u ? (unsigned)((gbase >> p) & 1u) : (egrp >> (p - 64u)) & 1u;
        sel |= b << j;
      }
    }
    off += sel << (flag ? 3 : 2);
  }
  return off;
}

// Register-bit mask of a lane flip word (0 where its selectors do not hold).
__device__ __forceinline__ unsigned flip_of(unsigned f, uint64_t gbase, unsigned egrp) {
  const unsigned p0 = (f >> 3) & 127u, p1 = (f >> 12) & 127u;
  const unsigned b0 = p0 < 64u ? (unsigned)((gbase >> p0) & 1u) : (egrp >> (p0 - 64u)) & 1u;
  bool ok = b0 == ((f >> 10) & 1u);
  if (f & 2048u) {
    const unsigned b1 = p1 < 64u ? (unsigned)((gbase >> p1) & 1u) : (egrp >> (p1 - 64u)) & 1u;
    ok = ok && b1 == ((f >> 19) & 1u);
  }
  return ok ? 1u << ((f >> 1) & 3u) : 0u;
}

// Controlled exchange with runtime control bits: the pair structure (target
// bit) is compile time, the control test a per-slot select, so no register
// is ever indexed at run time.  Slot r's pair swaps when its control bits
// (mask cm) equal cv.
template <int T>
__device__ __forceinline__ void sub_cx(v2d (&v)[kRegs], unsigned cm, unsigned cv) {
#pragma unroll
  for (int r = 0; r < kRegs; ++r) {
    if (r & (1 << T)) continue;
    const bool on = ((unsigned)r & cm) == cv;
    const v2d x = v[r], y = v[r | (1 << T)];
    v[r] = on ? y : x;
    v[r | (1 << T)] = on ? x : y;
  }
}

// Sub-op records are validated on the host (qfx_tile_check_ops), so every
// type and code below is one of the listed cases; the unreachable defaults
// let the dispatch drop its do-nothing path, which would otherwise copy all
// 16 register amplitudes on every sub-op.
__device__ __forceinline__ void apply_sub(v2d (&v)[kRegs], const int4 sub, const v2d* coef,
                                          uint64_t gbase, unsigned egrp, unsigned& fm) {
  const int type = sub.x, code = sub.y;
  const v2d* m = coef + sub.z;
  switch (type) {
    case SUB_D1: {
      // flagged: 8 coefficients, X m X in the second four for flipped lanes
      const v2d* mm = m + d1_offset(sub, fm, gbase, egrp);
      switch (code & 3) {
        case 0: sub_d1<0>(v, mm[0], mm[1], mm[2], mm[3]); break;
        case 1: sub_d1<1>(v, mm[0], mm[1], mm[2], mm[3]); break;
        case 2: sub_d1<2>(v, mm[0], mm[1], mm[2], mm[3]); break;
        case 3: sub_d1<3>(v, mm[0], mm[1], mm[2], mm[3]); break;
        default: __builtin_unreachable();
      }
      break;
    }
    case SUB_D1R: {
      const v2d* mm = m + d1_offset(sub, fm, gbase, egrp);
      switch (code & 3) {
        case 0: sub_d1r<0>(v, mm[0].x, mm[1].x, mm[2].x, mm[3].x); break;
        case 1: sub_d1r<1>(v, mm[0].x, mm[1].x, mm[2].x, mm[3].x); break;
        case 2: sub_d1r<2>(v, mm[0].x, mm[1].x, mm[2].x, mm[3].x); break;
        case 3: sub_d1r<3>(v, mm[0].x, mm[1].x, mm[2].x, mm[3].x); break;
        default: __builtin_unreachable();
      }
      break;
    }
    case SUB_D1X: {
      const v2d* mm = m + d1_offset(sub, fm, gbase, egrp);
      switch (code & 3) {
        case 0: sub_d1x<0>(v, mm[0].x, mm[1].y, mm[2].y, mm[3].x); break;
        case 1: sub_d1x<1>(v, mm[0].x, mm[1].y, mm[2].y, mm[3].x); break;
        case 2: sub_d1x<2>(v, mm[0].x, mm[1].y, mm[2].y, mm[3].x); break;
        case 3: sub_d1x<3>(v, mm[0].x, mm[1].y, mm[2].y, mm[3].x); break;
        default: __builtin_unreachable();
      }
      break;
    }
    case SUB_DIAG: sub_diag(v, m); break;
    case SUB_DIAGV: {
      // gbase: the tile's state index with its tile bits zero (uniform);
      // egrp: the lane group's tile index with its register bits zero
      const unsigned c = (unsigned)code;
      unsigned sel = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if (j < (int)(c & 3u)) {
          const unsigned p = (c >> (2 + 7 * j)) & 127u;
          const unsigned b = p < 64u ? (unsigned)((gbase >> p) & 1u) : (egrp >> (p - 64u)) & 1u;
          sel |= b << j;
        }
      }
      const unsigned w = (unsigned)sub.w;
      // flip-mask bits the tables read (w >> 8): slot r takes the entry of
      // slot r ^ (fm & those bits)
      const unsigned fx = (w >> 8) & 15u;
      if (w & 16u) {
        const unsigned dep = w & 15u;
        if (fx) sub_diag_compact_flip(v, m + (sel << __builtin_popcount(dep)), dep, fm & fx);
        else sub_diag_compact(v, m + (sel << __builtin_popcount(dep)), dep);
      } else {
        if (fx) sub_diag_flip(v, m + 16 * sel, fm & fx);
        else sub_diag(v, m + 16 * sel);
      }
      break;
    }
    case SUB_CX: {
      unsigned cm = ((unsigned)code >> 2) & 15u, cv = ((unsigned)code >> 6) & 15u;
      if (code & kCxFlip) {                 // control values as this lane holds them
        flagged_branch();
        cv ^= fm & cm;
      }
      const unsigned w = (unsigned)sub.w;
      if (w) {
        // selector controls: unmet, the test asks the target bit to be both
        // 0 (pair's low slot) and 1, so no pair swaps
        const bool ok = selectors_hold(w, gbase, egrp);
        const unsigned tb = 1u << (code & 3);
        cm = ok ? cm : (cm | tb);
        cv = ok ? cv : (cv | tb);
      }
      switch (code & 3) {
        case 0: sub_cx<0>(v, cm, cv); break;
        case 1: sub_cx<1>(v, cm, cv); break;
        case 2: sub_cx<2>(v, cm, cv); break;
        case 3: sub_cx<3>(v, cm, cv); break;
        default: __builtin_unreachable();
      }
      break;
    }
    default: __builtin_unreachable();
  }
}

// Uniform addressing of a register phase.  Logical index x of the phase
// (bit A <-> tile bit op.t[A]) sits at LDS slot sw(lb) ^ off(x), off XOR-
// linear; its affine load map (op.phase) and store map (op.active) give
// register r the logical index map(r) = c ^ (XOR of column j over the set
// bits j of r), so slot addresses are XOR trees over 4 basis values.  The
// basis values are the same for every tile: the host derives them once
// (phase_lds_words) into op.lds = {l[0..3], lc, s[0..3], sc}.
struct PhaseLds {
  unsigned l[4], lc, s[4], sc;
};

__host__ __device__ inline void phase_lds_words(const qfx_tile_op& op, uint32_t (&w)[10]) {
  unsigned b[4];
  for (int j = 0; j < 4; ++j) b[j] = sw(1u << op.t[j]);
  auto off = [&](unsigned x) {
    return ((x & 1u) ? b[0] : 0u) ^ ((x & 2u) ? b[1] : 0u) ^ ((x & 4u) ? b[2] : 0u) ^
           ((x & 8u) ? b[3] : 0u);
  };
  for (int j = 0; j < 4; ++j) {
    w[j] = off((op.phase >> (4 * j)) & 15u);
    w[5 + j] = off((op.active >> (4 * j)) & 15u);
  }
  w[4] = off(op.phase >> 16);
  w[9] = off(op.active >> 16);
}

// In bytes (slot x 16), so that an XOR tree yields the LDS byte address
// itself and no per-access scaling is needed.
__device__ __forceinline__ PhaseLds phase_lds(const qfx_tile_op& op) {
  PhaseLds p;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    p.l[j] = op.lds[j] << 4;
    p.s[j] = op.lds[5 + j] << 4;
  }
  p.lc = op.lds[4] << 4;
  p.sc = op.lds[9] << 4;
  return p;
}

template <typename U>
__device__ __forceinline__ void xor_tree(U (&a)[kRegs], U a0, const U (&b)[4]) {
  a[0] = a0;
  a[1] = a0 ^ b[0];
  a[2] = a0 ^ b[1];
  a[3] = a[1] ^ b[1];
#pragma unroll
  for (int r = 0; r < 4; ++r) a[4 + r] = a[r] ^ b[2];
#pragma unroll
  for (int r = 0; r < 8; ++r) a[8 + r] = a[r] ^ b[3];
}

// tile index of lane group g with zeros at the phase's register bits
__device__ __forceinline__ unsigned phase_lb(const qfx_tile_op& op, unsigned g) {
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) g = ins0(g, op.sorted[j]);
  return g;
}

// Returns the lane's flip mask, which the phase's store applies.
__device__ __forceinline__ unsigned run_subs(v2d (&v)[kRegs], const qfx_tile_op& op,
                                             const v2d* coef, const int* __restrict__ subs,
                                             uint64_t gbase, unsigned egrp) {
  // the next record's scalar load is issued with the current one's (scalar
  // loads complete out of order, so one wait covers both)
  const int4* rec = reinterpret_cast<const int4*>(subs + op.src);
  const int nsub = op.coef;
  unsigned fm = 0;
  int4 next = rec[0];
  // Two records per trip: every case of apply_sub writes its results to
  // fresh registers, and with one record per trip the loop latch copied
  // all 16 amplitudes back (32 64-bit moves per sub-op).  With two, the
  // second copy's results land in the loop's own registers.
  // A record's lane flips are applied outside the dispatch: a case that
  // left the amplitudes alone would bring the copies back.
  auto flips = [&](int4& r) {
    const unsigned f = (unsigned)r.x >> 8;
    if (f) fm ^= flip_of(f, gbase, egrp);
    r.x &= kRecType;
  };
  for (int i = 0; i < nsub; i += 2) {
    int4 cur = next;
    next = rec[i + 1 < nsub ? i + 1 : i];
    flips(cur);
    apply_sub(v, cur, coef, gbase, egrp, fm);
    if (i + 1 < nsub) {
      int4 cur2 = next;
      next = rec[i + 2 < nsub ? i + 2 : i + 1];
      flips(cur2);
      apply_sub(v, cur2, coef, gbase, egrp, fm);
    }
  }
  return fm;
}

// XOR of the basis values of the set bits of a flip mask: the address
// offset between register r and register r ^ fm.
template <typename U>
__device__ __forceinline__ U flip_off(unsigned fm, const U (&b)[4]) {
  return ((fm & 1u) ? b[0] : U(0)) ^ ((fm & 2u) ? b[1] : U(0)) ^ ((fm & 4u) ? b[2] : U(0)) ^
         ((fm & 8u) ? b[3] : U(0));
}

// The tile kernels declare no static LDS, so their dynamic LDS array starts
// at LDS address 0 (launch_pass checks this) and an XOR tree of byte offsets
// is the LDS address itself: no base add per access.
typedef __attribute__((address_space(3))) v2d v2d_lds;

__device__ __forceinline__ v2d_lds* lds_byte(unsigned a) {
  return (v2d_lds*)(size_t)a;
}

// slb: the lane group's base slot in bytes (sw(...) << 4)
__device__ __forceinline__ void phase_lds_load(v2d (&v)[kRegs], const v2d* /*lds*/, unsigned slb,
                                               const PhaseLds& p) {
  unsigned a[kRegs];
  xor_tree(a, slb ^ p.lc, p.l);
#pragma unroll
  for (int r = 0; r < kRegs; ++r) v[r] = *lds_byte(a[r]);
}

__device__ __forceinline__ void phase_lds_store(const v2d (&v)[kRegs], v2d* lds, unsigned slb,
                                                const PhaseLds& p, unsigned fm) {
  unsigned a[kRegs];
  xor_tree(a, slb ^ p.sc ^ flip_off(fm, p.s), p.s);
#pragma unroll
  for (int r = 0; r < kRegs; ++r) *lds_byte(a[r]) = v[r];
}

template <int BS>
__device__ __forceinline__ void op_rphase(v2d* lds, int T, const qfx_tile_op& op,
                                          const v2d* coef,
                                          const int* __restrict__ subs, uint64_t gbase) {
  const int ngroups = 1 << (T - kRegBits);
  const PhaseLds p = phase_lds(op);
  for (int g = threadIdx.x; g < ngroups; g += BS) {
    const unsigned egrp = phase_lb(op, (unsigned)g);
    const unsigned slb = sw(egrp) << 4;
    v2d v[kRegs];
    phase_lds_load(v, lds, slb, p);
    const unsigned fm = run_subs(v, op, coef, subs, gbase, egrp);
    phase_lds_store(v, lds, slb, p, fm);
  }
}

// ---- matrix-core blocks ---------------------------------------------------
// A dense k-qubit block Y = M X over all 2^(T-k) amplitude groups of a tile,
// computed as the real product [Re Y; Im Y] = R [Re X; Im X] with
// R = [[Re M, -Im M], [Im M, Re M]] on v_mfma_f64_16x16x4f64: A = a 16x4
// block of R, B = a 4x16 block of the real-form amplitudes (16 groups as
// columns), D = a 16x16 block of the result.  Operand layout of that
// instruction: lane l holds A[l%16][l/16], B[l/16][l%16] and, in register j
// of D, D[4j + l/16][l%16].  Lane l = (q = l/16, column c = l%16) therefore
// reads amplitudes 4s + q (s < 2^k/4) of group c, and the result rows it
// receives are the same amplitudes: each lane overwrites only what it read,
// and a column block belongs to one wave, so the op runs in place.
// R is packed on the host in operand order (P[(rb*NK + ks)*64 + l]).
typedef double v4d __attribute__((ext_vector_type(4)));

template <int K, int BS>
__device__ __forceinline__ void op_mfma(v2d* lds, int T, const qfx_tile_op& op,
                                        const v2d* __restrict__ coef) {
  constexpr int D = 1 << K;
  constexpr int NA = D / 4;       // amplitudes per lane and group
  constexpr int NK = 2 * D / 4;   // k-steps over the real form's columns
  constexpr int NR = 2 * D / 16;  // 16-row blocks of the real form
  const unsigned lane = threadIdx.x & 63u;
  const unsigned q = lane >> 4, c = lane & 15u;
  const double* __restrict__ P = reinterpret_cast<const double*>(coef + op.coef) + lane;
  unsigned soff[NA];
#pragma unroll
  for (int s = 0; s < NA; ++s) soff[s] = sw(loc_off<K>(op.t, 4u * s + q));
  const int ncb = 1 << (T - K - 4);   // column blocks of 16 groups
  for (int cb = threadIdx.x >> 6; cb < ncb; cb += BS / 64) {
    unsigned l = (unsigned)cb * 16u + c;
#pragma unroll
    for (int j = 0; j < K; ++j) l = ins0(l, op.sorted[j]);
    const unsigned sl = sw(l);
    v2d x[NA];
#pragma unroll
    for (int s = 0; s < NA; ++s) x[s] = lds[sl ^ soff[s]];
    // real-form row 4ks + q of the input: Re x[4ks + q] for ks < NA, else Im
    auto bval = [&](int ks) { return ks < NA ? x[ks].x : x[ks - NA].y; };
    if constexpr (K == 3) {
      // one row block: register j is row 4j + q = Re y[4j + q] (j < 2) or
      // Im y[4(j - 2) + q]
      v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < NK; ++ks)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(P[ks * 64], bval(ks), acc, 0, 0, 0);
      lds[sl ^ soff[0]] = v2d{acc[0], acc[2]};
      lds[sl ^ soff[1]] = v2d{acc[1], acc[3]};
    } else {
      // row blocks rp (real parts) and rp + NR/2 (imaginary parts) hold
      // amplitudes 4(4rp + j) + q in register j.  The 5-qubit block walks rp
      // at run time so its 4^6/64 packed-matrix values are re-read per
      // column block (L1 hits) instead of living in registers.
      constexpr int NP = NR / 2;
#pragma nounroll
      for (int rp = 0; rp < NP; ++rp) {
        v4d re = {0.0, 0.0, 0.0, 0.0}, im = {0.0, 0.0, 0.0, 0.0};
        const double* Pr = P + rp * NK * 64;
        const double* Pi = P + (rp + NP) * NK * 64;
#pragma unroll
        for (int ks = 0; ks < NK; ++ks) {
          const double b = bval(ks);
          re = __builtin_amdgcn_mfma_f64_16x16x4f64(Pr[ks * 64], b, re, 0, 0, 0);
          im = __builtin_amdgcn_mfma_f64_16x16x4f64(Pi[ks * 64], b, im, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned off = NP == 1 ? soff[j & (NA - 1)]
                                       : sw(loc_off<K>(op.t, 4u * (4u * rp + j) + q));
          lds[sl ^ off] = v2d{re[j], im[j]};
        }
      }
    }
  }
}

// `coef` holds the op list's coefficients (staged in LDS by the kernels that
// stage them); matrix-core blocks read their packed operands from `coef_g`,
// the same table in global memory (each lane loads its own operand).
template <int BS, bool WIDE, bool MF>
__device__ __forceinline__ void apply_op(v2d* lds, int T, const qfx_tile_op& op,
                                         const v2d* coef,
                                         const v2d* __restrict__ coef_g,
                                         const int* __restrict__ srcs, uint64_t gbase) {
  if (op.kind == OP_RPHASE) {
    // 16 register amplitudes per lane exceed what 1024-lane blocks leave
    if constexpr (BS <= kRphaseMaxBlock) op_rphase<BS>(lds, T, op, coef, srcs, gbase);
  } else if (op.kind == OP_MFMA) {
    if constexpr (MF && BS <= kMfmaMaxBlock) {
      switch (op.k) {
        case 3: op_mfma<3, BS>(lds, T, op, coef_g); break;
        case 4: op_mfma<4, BS>(lds, T, op, coef_g); break;
        case 5: op_mfma<5, BS>(lds, T, op, coef_g); break;
      }
    }
  } else if (op.kind == OP_DENSE) {
    switch (op.k) {
      case 1: op_dense<1, BS>(lds, T, op, coef); break;
      case 2: op_dense<2, BS>(lds, T, op, coef); break;
      case 3: op_dense<3, BS>(lds, T, op, coef); break;
      // 4- and 5-qubit ops need more registers than blocks of more than
      // 256 lanes leave a lane; the host validator rejects them there
      case 4:
        if constexpr (WIDE && BS <= kWideOpMaxBlock) op_dense<4, BS>(lds, T, op, coef);
        break;
      case 5:
        if constexpr (WIDE && BS <= kWideOpMaxBlock) op_dense<5, BS>(lds, T, op, coef);
        break;
    }
  } else {
    switch (op.k) {
      case 1: op_sparse<1, BS>(lds, T, op, coef, srcs); break;
      case 2: op_sparse<2, BS>(lds, T, op, coef, srcs); break;
      case 3: op_sparse<3, BS>(lds, T, op, coef, srcs); break;
      case 4:
        if constexpr (WIDE && BS <= kWideOpMaxBlock) op_sparse<4, BS>(lds, T, op, coef, srcs);
        break;
      case 5:
        if constexpr (WIDE && BS <= kWideOpMaxBlock) op_sparse<5, BS>(lds, T, op, coef, srcs);
        break;
    }
  }
}

// Global offset of the tile bits above the lane bits: bit j of kk sits at
// index bit pos.s[LB + j].
template <int LB>
__device__ __forceinline__ uint64_t high_off(int kk, int T, const int* bitpos) {
  uint64_t o = 0;
#pragma unroll
  for (int j = 0; j < kMaxTileBits - LB; ++j)
    if (j < T - LB) o |= (uint64_t)((kk >> j) & 1) << bitpos[LB + j];
  return o;
}

template <int BS>
__device__ __forceinline__ v2d* stage_coef(v2d* lds, int T, const v2d* __restrict__ coef, int ncoef) {
  v2d* cf = lds + (1 << T);
  for (int i = threadIdx.x; i < ncoef; i += BS) cf[i] = coef[i];
  return cf;   // read after the kernel's first barrier
}

template <int BS, int LB, bool WIDE, bool MF>
__global__ __launch_bounds__(BS) void tile_pass_kernel(v2d* state, int T, TilePos pos,
                                                      const qfx_tile_op* __restrict__ ops,
                                                      int nops, const v2d* __restrict__ coef,
                                                      int ncoef, const int* __restrict__ srcs) {
  extern __shared__ v2d lds[];
  const int tid = threadIdx.x;
  constexpr bool kStage = !WIDE && !MF;
  v2d* cf = kStage ? stage_coef<BS>(lds, T, coef, ncoef) : nullptr;
  // tile base: block index bits spread over the non-tile positions
  // (loops over tile bits run to a compile-time bound so pos.s is indexed
  // with constants and stays in kernel-argument registers)
  uint64_t base = blockIdx.x;
#pragma unroll
  for (int j = 0; j < kMaxTileBits; ++j) {
    if (j < T) {
      const int b = pos.s[j];
      base = ((base >> b) << (b + 1)) | (base & (((uint64_t)1 << b) - 1));
    }
  }
  uint64_t toff = 0, soff = 0;
#pragma unroll
  for (int j = 0; j < LB; ++j) {
    toff |= (uint64_t)((tid >> j) & 1) << pos.s[j];
    soff |= (uint64_t)((tid >> j) & 1) << pos.d[j];
  }
  const int per = 1 << (T - LB);   // amplitudes per lane
  const v2d* src = state + base + toff;
  // issue kLoadBatch global loads before the first LDS write
  for (int k0 = 0; k0 < per; k0 += kLoadBatch) {
    v2d v[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int kk = k0 + u;
      if (kk < per) {
        v[u] = ld_stream(src + high_off<LB>(kk, T, pos.s));
      }
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u)
      if (k0 + u < per) lds[sw(tid + (k0 + u) * BS)] = v[u];
  }
  __syncthreads();
  for (int i = 0; i < nops; ++i) {
    if constexpr (kStage) apply_op<BS, WIDE, MF>(lds, T, ops[i], cf, coef, srcs, base);
    else apply_op<BS, WIDE, MF>(lds, T, ops[i], coef, coef, srcs, base);
    __syncthreads();
  }
  v2d* dst = state + base + soff;
  for (int k0 = 0; k0 < per; k0 += kLoadBatch) {
    v2d v[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u)
      if (k0 + u < per) v[u] = lds[sw(tid + (k0 + u) * BS)];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int kk = k0 + u;
      if (kk < per) {
        st_stream(dst + high_off<LB>(kk, T, pos.d), v[u]);
      }
    }
  }
}


// Persistent, software-pipelined variant: each workgroup walks tiles
// blockIdx.x, blockIdx.x + gridDim.x, ...; while the current tile is
// processed in LDS, the next tile's amplitudes are already in flight into
// registers (pf), so HBM traffic overlaps the LDS work.  Needs exactly
// kPrefetch = 2^T / BS amplitudes per lane (T = LB + 4): with that count
// fixed at compile time the loads and stores carry no guards, and the wait
// before a prefetched amplitude goes to LDS covers only its own load, not
// the previous tile's stores.
constexpr int kPrefetch = 16;

template <int BS, int LB, bool WIDE, bool MF>
__global__ __launch_bounds__(BS) void tile_pipe_kernel(v2d* state, int T, TilePos pos,
                                                      const qfx_tile_op* __restrict__ ops,
                                                      int nops, const v2d* __restrict__ coef,
                                                      int ncoef, const int* __restrict__ srcs,
                                                      uint32_t ntiles) {
  extern __shared__ v2d lds[];
  const int tid = threadIdx.x;
  constexpr bool kStage = !WIDE && !MF;
  v2d* cf = kStage ? stage_coef<BS>(lds, T, coef, ncoef) : nullptr;
  uint64_t toff = 0, soff = 0;
#pragma unroll
  for (int j = 0; j < LB; ++j) {
    toff |= (uint64_t)((tid >> j) & 1) << pos.s[j];
    soff |= (uint64_t)((tid >> j) & 1) << pos.d[j];
  }
  // the four tile bits above the lane bits: amplitude u of a lane sits at
  // h(u) = sum of hb[j] over the set bits j of u (stores: pos.d, sb[j])
  const uint64_t hb0 = (uint64_t)1 << pos.s[LB], hb1 = (uint64_t)1 << pos.s[LB + 1];
  const uint64_t hb2 = (uint64_t)1 << pos.s[LB + 2], hb3 = (uint64_t)1 << pos.s[LB + 3];
  const uint64_t sb0 = (uint64_t)1 << pos.d[LB], sb1 = (uint64_t)1 << pos.d[LB + 1];
  const uint64_t sb2 = (uint64_t)1 << pos.d[LB + 2], sb3 = (uint64_t)1 << pos.d[LB + 3];

  auto tile_base = [&](uint32_t tile) {
    uint64_t base = tile;
#pragma unroll
    for (int j = 0; j < kMaxTileBits; ++j) {
      if (j < T) {
        const int b = pos.s[j];
        base = ((base >> b) << (b + 1)) | (base & (((uint64_t)1 << b) - 1));
      }
    }
    return base;
  };
  auto load_tile = [&](v2d (&pf)[kPrefetch], uint64_t base) {
    const v2d* p0 = state + base + toff;
    const v2d* p1 = p0 + hb1;
#pragma unroll
    for (int u = 0; u < kPrefetch; u += 4) {
      const uint64_t o = ((u & 4) ? hb2 : 0) + ((u & 8) ? hb3 : 0);
      pf[u] = ld_stream(p0 + o);
      pf[u + 1] = ld_stream(p0 + o + hb0);
      pf[u + 2] = ld_stream(p1 + o);
      pf[u + 3] = ld_stream(p1 + o + hb0);
    }
  };

  // the tile number's bits are deposited at the non-tile index bits; adding
  // the deposited stride with the tile bits set to one carries across them
  uint64_t tmask = 0;
#pragma unroll
  for (int j = 0; j < kMaxTileBits; ++j)
    if (j < T) tmask |= (uint64_t)1 << pos.s[j];
  const uint64_t step = tile_base(gridDim.x);
  v2d pf[kPrefetch];
  uint32_t tile = blockIdx.x;       // the grid never exceeds ntiles
  uint64_t base = tile_base(tile);
  load_tile(pf, base);
#pragma unroll
  for (int u = 0; u < kPrefetch; ++u) lds[sw(tid + u * BS)] = pf[u];
  for (; tile < ntiles; tile += gridDim.x) {
    // the last tile re-reads itself instead of branching around the loads
    const uint64_t nbase =
        tile + gridDim.x < ntiles ? ((base | tmask) + step) & ~tmask : base;
    load_tile(pf, nbase);
    __syncthreads();
    for (int i = 0; i < nops; ++i) {
      if constexpr (kStage) apply_op<BS, WIDE, MF>(lds, T, ops[i], cf, coef, srcs, base);
      else apply_op<BS, WIDE, MF>(lds, T, ops[i], coef, coef, srcs, base);
      __syncthreads();
    }
    v2d* q0 = state + base + soff;
    v2d* q1 = q0 + sb1;
#pragma unroll
    for (int u = 0; u < kPrefetch; u += 4) {
      const uint64_t o = ((u & 4) ? sb2 : 0) + ((u & 8) ? sb3 : 0);
      st_stream(q0 + o, lds[sw(tid + u * BS)]);
      st_stream(q0 + o + sb0, lds[sw(tid + (u + 1) * BS)]);
      st_stream(q1 + o, lds[sw(tid + (u + 2) * BS)]);
      st_stream(q1 + o + sb0, lds[sw(tid + (u + 3) * BS)]);
    }
    base = nbase;
    __syncthreads();
    // the next tile goes to LDS here, in the iteration that loaded it: the
    // wait for its loads then counts the stores issued after them (vector
    // memory operations complete in order) instead of draining them
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) lds[sw(tid + u * BS)] = pf[u];
  }
}

// ---- register-phase passes that store through their last phase ----------
// A persistent pass (one 16-amplitude group per lane, T = LB + 4) whose
// last op is a register phase stores that phase's registers straight to
// HBM, saving the LDS write of its results, the LDS read of the store loop
// and a barrier per tile.  Global offsets: tile bit j <-> index bit
// pos.d[j]; a logical index x of the phase sits at the sum of 2^d[t[A]]
// over its set bits A, XOR-linear as the LDS offsets.
struct PhaseGlobal {
  uint64_t b[4], c;
};

__device__ __forceinline__ uint64_t tile_bit_off(int tbit, int T, const int* bitpos) {
  uint64_t o = 0;
#pragma unroll
  for (int j = 0; j < kMaxTileBits; ++j)
    if (j < T && j == tbit) o = (uint64_t)1 << bitpos[j];
  return o;
}

__device__ __forceinline__ uint64_t tile_state_off(unsigned e, int T, const int* bitpos) {
  uint64_t o = 0;
#pragma unroll
  for (int j = 0; j < kMaxTileBits; ++j)
    if (j < T) o |= (uint64_t)((e >> j) & 1u) << bitpos[j];
  return o;
}

__device__ __forceinline__ PhaseGlobal phase_global(const qfx_tile_op& op, unsigned map, int T,
                                                    const int* bitpos) {
  uint64_t e[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) e[a] = tile_bit_off(op.t[a], T, bitpos);
  auto off = [&](unsigned x) {
    return ((x & 1u) ? e[0] : 0) ^ ((x & 2u) ? e[1] : 0) ^ ((x & 4u) ? e[2] : 0) ^
           ((x & 8u) ? e[3] : 0);
  };
  PhaseGlobal g;
#pragma unroll
  for (int j = 0; j < 4; ++j) g.b[j] = off((map >> (4 * j)) & 15u);
  g.c = off(map >> 16);
  return g;
}

__device__ __forceinline__ void phase_gstore(const v2d (&v)[kRegs], v2d* p, const PhaseGlobal& g,
                                             unsigned fm) {
  uint64_t a[kRegs];
  xor_tree(a, g.c ^ flip_off(fm, g.b), g.b);
#pragma unroll
  for (int r = 0; r < kRegs; ++r) st_stream(p + a[r], v[r]);
}

template <int BS, int LB>
__global__ __launch_bounds__(BS) void tile_reg_kernel(v2d* state, int T, TilePos pos,
                                                     const qfx_tile_op* __restrict__ ops,
                                                     int nops, const v2d* __restrict__ coef,
                                                     int ncoef, const int* __restrict__ srcs,
                                                     uint32_t ntiles) {
  extern __shared__ v2d lds[];
  const int tid = threadIdx.x;
  v2d* cf = stage_coef<BS>(lds, T, coef, ncoef);
  const qfx_tile_op& last = ops[nops - 1];
  // the lane's group in the last phase, as HBM offsets
  const uint64_t l_lane = tile_state_off(phase_lb(last, (unsigned)tid), T, pos.d);
  const PhaseGlobal lg = phase_global(last, last.active, T, pos.d);
  uint64_t toff = 0;
#pragma unroll
  for (int j = 0; j < LB; ++j) toff |= (uint64_t)((tid >> j) & 1) << pos.s[j];
  const uint64_t hb0 = (uint64_t)1 << pos.s[LB], hb1 = (uint64_t)1 << pos.s[LB + 1];
  const uint64_t hb2 = (uint64_t)1 << pos.s[LB + 2], hb3 = (uint64_t)1 << pos.s[LB + 3];

  auto tile_base = [&](uint32_t tile) {
    uint64_t base = tile;
#pragma unroll
    for (int j = 0; j < kMaxTileBits; ++j) {
      if (j < T) {
        const int b = pos.s[j];
        base = ((base >> b) << (b + 1)) | (base & (((uint64_t)1 << b) - 1));
      }
    }
    return base;
  };
  auto load_tile = [&](v2d (&pf)[kPrefetch], uint64_t base) {
    const v2d* p0 = state + base + toff;
    const v2d* p1 = p0 + hb1;
#pragma unroll
    for (int u = 0; u < kPrefetch; u += 4) {
      const uint64_t o = ((u & 4) ? hb2 : 0) + ((u & 8) ? hb3 : 0);
      pf[u] = ld_stream(p0 + o);
      pf[u + 1] = ld_stream(p0 + o + hb0);
      pf[u + 2] = ld_stream(p1 + o);
      pf[u + 3] = ld_stream(p1 + o + hb0);
    }
  };

  // the tile number's bits are deposited at the non-tile index bits; adding
  // the deposited stride with the tile bits set to one carries across them
  uint64_t tmask = 0;
#pragma unroll
  for (int j = 0; j < kMaxTileBits; ++j)
    if (j < T) tmask |= (uint64_t)1 << pos.s[j];
  const uint64_t step = tile_base(gridDim.x);
  v2d pf[kPrefetch];
  uint32_t tile = blockIdx.x;       // the grid never exceeds ntiles
  uint64_t base = tile_base(tile);
  load_tile(pf, base);
#pragma unroll
  for (int u = 0; u < kPrefetch; ++u) lds[sw(tid + u * BS)] = pf[u];
  for (; tile < ntiles; tile += gridDim.x) {
    // the last tile re-reads itself instead of branching around the loads
    const uint64_t nbase =
        tile + gridDim.x < ntiles ? ((base | tmask) + step) & ~tmask : base;
    load_tile(pf, nbase);
    __syncthreads();
    for (int i = 0; i < nops - 1; ++i) {
      apply_op<BS, false, false>(lds, T, ops[i], cf, coef, srcs, base);
      __syncthreads();
    }
    {
      const PhaseLds lp = phase_lds(last);
      v2d v[kRegs];
      const unsigned egrp = phase_lb(last, (unsigned)tid);
      phase_lds_load(v, lds, sw(egrp) << 4, lp);
      const unsigned fm = run_subs(v, last, cf, srcs, base, egrp);
      phase_gstore(v, state + base + l_lane, lg, fm);
    }
    base = nbase;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) lds[sw(tid + u * BS)] = pf[u];
  }
}

// ---- register-phase passes that load through their first phase ----------
// A persistent pass (T = LB + 4) whose first op is a register phase loads
// that phase's 16 amplitudes per lane straight from HBM into registers (the
// phase's load map applied to global offsets, tile bit j <-> index bit
// pos.s[j]), so the tile never takes the LDS write, barrier and LDS read
// that would otherwise precede the phase.  The next tile's amplitudes are
// in flight in `pf` while the current one is processed.  With RIO the last
// register phase also stores straight to HBM; a pass of one such phase
// then touches no LDS at all.  Lanes 0..7 of a wave own neighbouring
// amplitudes when the phase's register bits avoid tile bits 0..2 (the
// planner asks for that), so loads come in 128-byte runs.
__device__ __forceinline__ void phase_gload(v2d (&v)[kRegs], const v2d* p, const PhaseGlobal& g) {
  uint64_t a[kRegs];
  xor_tree(a, g.c, g.b);
#pragma unroll
  for (int r = 0; r < kRegs; ++r) v[r] = ld_stream(p + a[r]);
}

// How a register-input pass ends: its tile goes back to HBM from LDS
// (RIN_LDS_OUT), through its last register phase (RIN_REG_OUT), or the pass
// is that one phase, loading and storing through registers (RIN_SINGLE).
enum RinMode : int { RIN_LDS_OUT = 0, RIN_REG_OUT = 1, RIN_SINGLE = 2 };

// Two waves per SIMD (at most 256 registers a lane, the other kernels'
// budget) is what hides the LDS and memory latency; the hint keeps the
// variants that hold the next tile and a phase's registers at once within it.
template <int BS, int LB, int MODE>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2))) void tile_rin_kernel(
    v2d* state, int T, TilePos pos, const qfx_tile_op* __restrict__ ops, int nops,
    const v2d* __restrict__ coef, int ncoef, const int* __restrict__ srcs, uint32_t ntiles) {
  extern __shared__ v2d lds[];
  const int tid = threadIdx.x;
  v2d* cf = stage_coef<BS>(lds, T, coef, ncoef);
  const qfx_tile_op& first = ops[0];
  const unsigned f_egrp = phase_lb(first, (unsigned)tid);
  const uint64_t f_lane = tile_state_off(f_egrp, T, pos.s);
  const PhaseGlobal fg = phase_global(first, first.phase, T, pos.s);
  uint64_t tmask = 0;
#pragma unroll
  for (int j = 0; j < kMaxTileBits; ++j)
    if (j < T) tmask |= (uint64_t)1 << pos.s[j];
  auto tile_base = [&](uint32_t t) {
    uint64_t base = t;
#pragma unroll
    for (int j = 0; j < kMaxTileBits; ++j) {
      if (j < T) {
        const int b = pos.s[j];
        base = ((base >> b) << (b + 1)) | (base & (((uint64_t)1 << b) - 1));
      }
    }
    return base;
  };
  const uint64_t step = tile_base(gridDim.x);
  v2d pf[kRegs];
  uint32_t tile = blockIdx.x;       // the grid never exceeds ntiles
  uint64_t base = tile_base(tile);
  phase_gload(pf, state + base + f_lane, fg);
  __syncthreads();                  // staged coefficients
  if constexpr (MODE == RIN_SINGLE) {
    const uint64_t s_lane = tile_state_off(f_egrp, T, pos.d);
    const PhaseGlobal sg = phase_global(first, first.active, T, pos.d);
    for (; tile < ntiles; tile += gridDim.x) {
      const uint64_t nbase =
          tile + gridDim.x < ntiles ? ((base | tmask) + step) & ~tmask : base;
      v2d v[kRegs];
#pragma unroll
      for (int r = 0; r < kRegs; ++r) v[r] = pf[r];
      phase_gload(pf, state + nbase + f_lane, fg);
      const unsigned fm = run_subs(v, first, cf, srcs, base, f_egrp);
      // No LDS is involved, but with a permuted store a wave writes
      // amplitudes other waves of the workgroup load: every wave must hold
      // this tile in registers (v) before any stores it.
      if (pos.permuted) __syncthreads();
      phase_gstore(v, state + base + s_lane, sg, fm);
      base = nbase;
    }
    return;
  }
  // store side of RIN_LDS_OUT: the LDS store loop
  uint64_t soff = 0;
#pragma unroll
  for (int j = 0; j < LB; ++j) soff |= (uint64_t)((tid >> j) & 1) << pos.d[j];
  const uint64_t sb0 = (uint64_t)1 << pos.d[LB], sb1 = (uint64_t)1 << pos.d[LB + 1];
  const uint64_t sb2 = (uint64_t)1 << pos.d[LB + 2], sb3 = (uint64_t)1 << pos.d[LB + 3];
  const int nmid_end = MODE == RIN_REG_OUT ? nops - 1 : nops;   // ops 1 .. nmid_end-1 from LDS
  for (; tile < ntiles; tile += gridDim.x) {
    // the last tile re-reads itself instead of branching around the loads
    const uint64_t nbase =
        tile + gridDim.x < ntiles ? ((base | tmask) + step) & ~tmask : base;
    v2d v[kRegs];
#pragma unroll
    for (int r = 0; r < kRegs; ++r) v[r] = pf[r];
    phase_gload(pf, state + nbase + f_lane, fg);
    const unsigned ffm = run_subs(v, first, cf, srcs, base, f_egrp);
    phase_lds_store(v, lds, sw(f_egrp) << 4, phase_lds(first), ffm);
    __syncthreads();
    for (int i = 1; i < nmid_end; ++i) {
      apply_op<BS, false, false>(lds, T, ops[i], cf, coef, srcs, base);
      __syncthreads();
    }
    if constexpr (MODE == RIN_REG_OUT) {
      // the last phase reuses the first phase's registers
      const qfx_tile_op& last = ops[nops - 1];
      const unsigned egrp = phase_lb(last, (unsigned)tid);
      phase_lds_load(v, lds, sw(egrp) << 4, phase_lds(last));
      const unsigned lfm = run_subs(v, last, cf, srcs, base, egrp);
      phase_gstore(v, state + base + tile_state_off(egrp, T, pos.d),
                   phase_global(last, last.active, T, pos.d), lfm);
    } else {
      v2d* q0 = state + base + soff;
      v2d* q1 = q0 + sb1;
#pragma unroll
      for (int u = 0; u < kPrefetch; u += 4) {
        const uint64_t o = ((u & 4) ? sb2 : 0) + ((u & 8) ? sb3 : 0);
        st_stream(q0 + o, lds[sw(tid + u * BS)]);
        st_stream(q0 + o + sb0, lds[sw(tid + (u + 1) * BS)]);
        st_stream(q1 + o, lds[sw(tid + (u + 2) * BS)]);
        st_stream(q1 + o + sb0, lds[sw(tid + (u + 3) * BS)]);
      }
    }
    base = nbase;
    __syncthreads();                // LDS free for the next tile's first phase
  }
}

// Register phases address LDS by absolute byte offsets (lds_byte), valid
// only when the kernel's dynamic LDS array starts at address 0, i.e. the
// kernel has no static LDS; also raises the dynamic LDS limit.
int prepare_kernel(const void* fn, const char* name) {
  hipFuncAttributes attr;
  int rc = qfx::check_hip(hipFuncGetAttributes(&attr, fn), "hipFuncGetAttributes");
  if (rc) return rc;
  if (attr.sharedSizeBytes != 0)
    return qfx::set_error(QFX_EKERNEL, std::string(name) + ": unexpected static LDS");
  return qfx::check_hip(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            160 * 1024),
                        "hipFuncSetAttribute");
}

// Per (kernel, device): the kernel's attributes are set once, and its
// persistent grid is the workgroups of `block` lanes with `lds` bytes that
// fit the device at once (occupancy x CUs), cached per LDS size.
int kernel_grid(const void* fn, const char* name, int block, size_t lds, int* grid) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, int, size_t>, int> grids;
  static std::set<std::pair<const void*, int>> prepared;
  int dev = 0;
  int rc = qfx::check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(mu);
  if (!prepared.count({fn, dev})) {
    rc = prepare_kernel(fn, name);
    if (rc) return rc;
    prepared.insert({fn, dev});
  }
  if (!grid) return QFX_OK;
  auto key = std::make_tuple(fn, dev, lds);
  auto it = grids.find(key);
  if (it == grids.end()) {
    int per_cu = 0, cus = 0;
    rc = qfx::check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds),
                        "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (rc) return rc;
    rc = qfx::check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                        "hipDeviceGetAttribute");
    if (rc) return rc;
    it = grids.emplace(key, (per_cu > 0 ? per_cu : 1) * cus).first;
  }
  *grid = it->second;
  return QFX_OK;
}

template <int BS, int LB, bool WIDE, bool MF>
int launch_pass(hipStream_t s, v2d* state, int n, int T, const TilePos& pos,
                const qfx_tile_op* ops, int nops, const v2d* coef, int ncoef, const int* srcs,
                bool persistent, bool reg_io, bool reg_in) {
  constexpr bool kStage = !WIDE && !MF;
  const size_t lds = sizeof(v2d) * ((size_t(1) << T) + (kStage ? ncoef : 0));
  // occupancy (persistent grid size) for the largest staged table
  const size_t lds_max = sizeof(v2d) * ((size_t(1) << T) + (kStage ? kMaxStaged : 0));
  const uint64_t tiles = (uint64_t)1 << (n - T);
  int grid = 0;
  int rc = QFX_OK;
  if constexpr (!WIDE && !MF && BS <= kRphaseMaxBlock) {
    // the caller checked T == LB + 4 for these
    if (reg_in) {
      const int mode = !reg_io ? RIN_LDS_OUT : (nops == 1 ? RIN_SINGLE : RIN_REG_OUT);
      const void* fn = mode == RIN_LDS_OUT   ? (const void*)tile_rin_kernel<BS, LB, RIN_LDS_OUT>
                       : mode == RIN_SINGLE ? (const void*)tile_rin_kernel<BS, LB, RIN_SINGLE>
                                            : (const void*)tile_rin_kernel<BS, LB, RIN_REG_OUT>;
      rc = kernel_grid(fn, "tile_rin_kernel", BS, lds_max, &grid);
      if (rc) return rc;
      const uint64_t g = tiles < (uint64_t)grid ? tiles : (uint64_t)grid;
      const uint32_t nt = (uint32_t)tiles;
      if (mode == RIN_LDS_OUT)
        hipLaunchKernelGGL((tile_rin_kernel<BS, LB, RIN_LDS_OUT>), dim3((unsigned)g), dim3(BS),
                           lds, s, state, T, pos, ops, nops, coef, ncoef, srcs, nt);
      else if (mode == RIN_SINGLE)
        hipLaunchKernelGGL((tile_rin_kernel<BS, LB, RIN_SINGLE>), dim3((unsigned)g), dim3(BS),
                           lds, s, state, T, pos, ops, nops, coef, ncoef, srcs, nt);
      else
        hipLaunchKernelGGL((tile_rin_kernel<BS, LB, RIN_REG_OUT>), dim3((unsigned)g), dim3(BS),
                           lds, s, state, T, pos, ops, nops, coef, ncoef, srcs, nt);
      return qfx::check_launch("tile_rin_kernel");
    }
    if (reg_io) {
      rc = kernel_grid((const void*)tile_reg_kernel<BS, LB>, "tile_reg_kernel", BS, lds_max, &grid);
      if (rc) return rc;
      const uint64_t g = tiles < (uint64_t)grid ? tiles : (uint64_t)grid;
      hipLaunchKernelGGL((tile_reg_kernel<BS, LB>), dim3((unsigned)g), dim3(BS), lds, s, state,
                         T, pos, ops, nops, coef, ncoef, srcs, (uint32_t)tiles);
      return qfx::check_launch("tile_reg_kernel");
    }
  }
  if (persistent && BS <= 512 && T - LB == 4) {
    rc = kernel_grid((const void*)tile_pipe_kernel<BS, LB, WIDE, MF>, "tile_pipe_kernel", BS,
                     lds_max, &grid);
    if (rc) return rc;
    const uint64_t g = tiles < (uint64_t)grid ? tiles : (uint64_t)grid;
    hipLaunchKernelGGL((tile_pipe_kernel<BS, LB, WIDE, MF>), dim3((unsigned)g), dim3(BS), lds, s,
                       state, T, pos, ops, nops, coef, ncoef, srcs, (uint32_t)tiles);
    return qfx::check_launch("tile_pipe_kernel");
  }
  rc = kernel_grid((const void*)tile_pass_kernel<BS, LB, WIDE, MF>, "tile_pass_kernel", BS, lds,
                   nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL((tile_pass_kernel<BS, LB, WIDE, MF>), dim3((unsigned)tiles), dim3(BS), lds, s, state, T,
                     pos, ops, nops, coef, ncoef, srcs);
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

int qfx_tile_prepare_ops(qfx_tile_op* ops, int nops) {
  if (nops < 0 || (!ops && nops > 0))
    return qfx::set_error(QFX_EINVAL, "qfx_tile_prepare_ops: bad op list");
  for (int i = 0; i < nops; ++i) {
    qfx_tile_op& o = ops[i];
    uint32_t w[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (o.kind == OP_RPHASE) {
      for (int j = 0; j < kRegBits; ++j)
        if (o.t[j] < 0 || o.t[j] >= kMaxTileBits)
          return qfx::set_error(QFX_EINVAL, "qfx_tile_prepare_ops: register bit outside a tile");
      phase_lds_words(o, w);
    }
    for (int j = 0; j < 10; ++j) o.lds[j] = w[j];
  }
  return QFX_OK;
}

int qfx_tile_check_ops(int T, int block, int flags, const qfx_tile_op* ops, int nops,
                       int ncoef, const int* srcs, int nsrcs) {
  if (flags & ~(QFX_TILE_WIDE | QFX_TILE_PERSISTENT | QFX_TILE_MFMA | QFX_TILE_REG_IO |
                QFX_TILE_REG_IN))
    return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: unknown flags");
  if (flags & QFX_TILE_REG_IN) {
    if ((flags & (QFX_TILE_WIDE | QFX_TILE_MFMA)) || block > kRphaseMaxBlock ||
        (1 << T) != block * kRegs || nops < 1 || !ops || ops[0].kind != OP_RPHASE)
      return qfx::set_error(QFX_EINVAL,
                            "qfx_tile_check_ops: QFX_TILE_REG_IN needs a first register phase, "
                            "16 amplitudes per lane, block <= 512");
  }
  if (flags & QFX_TILE_REG_IO) {
    if ((flags & (QFX_TILE_WIDE | QFX_TILE_MFMA)) || block > kRphaseMaxBlock ||
        (1 << T) != block * kRegs || nops < 1 || !ops || ops[nops - 1].kind != OP_RPHASE)
      return qfx::set_error(QFX_EINVAL,
                            "qfx_tile_check_ops: QFX_TILE_REG_IO needs a last register phase, "
                            "16 amplitudes per lane, block <= 512");
  }
  if ((flags & QFX_TILE_WIDE) && block > kWideOpMaxBlock)
    return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: 4-5 qubit ops need block <= 256");
  if ((flags & QFX_TILE_MFMA) && block > kMfmaMaxBlock)
    return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: matrix-core blocks need block <= 512");
  if (!ops && nops > 0) return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: null ops");
  if (!(flags & (QFX_TILE_WIDE | QFX_TILE_MFMA)) && ncoef > kMaxStaged)
    return qfx::set_error(QFX_EINVAL,
                          "qfx_tile_check_ops: more than 768 coefficients in a pass without wide ops");
  const int kmax = (flags & QFX_TILE_WIDE) ? 5 : 3;
  for (int i = 0; i < nops; ++i) {
    const qfx_tile_op& o = ops[i];
    if (o.kind == OP_MFMA) {
      if (!(flags & QFX_TILE_MFMA))
        return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: matrix-core block without QFX_TILE_MFMA");
      if (o.k < 3 || o.k > 5 || T < o.k + 4)
        return qfx::set_error(QFX_EINVAL,
                              "qfx_tile_check_ops: matrix-core blocks need 3 <= k <= 5 and T >= k + 4");
      if (o.coef < 0 || o.coef + 2 * (1 << (2 * o.k)) > ncoef)
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
      continue;
    }
    if (o.kind == OP_RPHASE) {
      if (T < kRegBits) return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: tile too small");
      if (block > kRphaseMaxBlock)
        return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: register phases need block <= 512");
      for (int j = 0; j < kRegBits; ++j) {
        if (o.t[j] < 0 || o.t[j] >= T)
          return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: register bit outside the tile");
        if (j > 0 && o.sorted[j] <= o.sorted[j - 1])
          return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: register bits not distinct");
        bool found = false;
        for (int i2 = 0; i2 < kRegBits; ++i2) found |= o.sorted[i2] == o.t[j];
        if (!found) return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: sorted != register bits");
      }
      // load and store maps: affine, invertible (4 independent columns)
      for (const uint32_t m : {o.phase, o.active}) {
        unsigned span = 1u;                   // bit x set: x is in the column span
        for (int j = 0; j < kRegBits; ++j) {
          const unsigned col = (m >> (4 * j)) & 15u;
          unsigned next = span;
          for (unsigned x = 0; x < 16; ++x)
            if ((span >> x) & 1u) next |= 1u << (x ^ col);
          span = next;
        }
        if ((m >> 20) != 0 || span != 0xFFFFu)
          return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: bad register-phase load/store map");
      }
      {
        uint32_t w[10];
        phase_lds_words(o, w);
        for (int j = 0; j < 10; ++j)
          if (o.lds[j] != w[j])
            return qfx::set_error(QFX_EINVAL,
                                  "qfx_tile_check_ops: register-phase LDS words do not match "
                                  "its bits and maps (run qfx_tile_prepare_ops)");
      }
      // the sub-op loop reads the record at src even when the list is empty
      if (o.coef < 0 || o.src < 0 || o.src % 4 != 0 ||
          o.src + 4 * (o.coef > 0 ? o.coef : 1) > nsrcs)
        return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: sub-ops out of range");
      for (int q = 0; q < o.coef; ++q) {
        const int* sub = srcs + o.src + 4 * q;
        // codes per type: D1/D1R/D1X 4 register bits (+ kD1Flip: 8
        // coefficients), DIAG 0, CX 4 targets x 16 control masks x 16
        // control values (+ kCxFlip); control masks may not touch the target
        // DIAGV: 1..3 selector bits, each an index bit < 64, 16 << m coefs
        static const int ncode[9] = {0, 8, 0, 0, 1, 2048, 0, 8, 8};
        static const int need[9] = {0, 4, 0, 0, kRegs, 0, 0, 4, 4};
        // a selector code: an index bit < 64 outside the tile (the caller's
        // tile positions are not known here), or 64 + a tile bit outside the
        // registers
        auto selector_ok = [&](unsigned p) {
          if (p < 64u) return true;
          const int t = (int)(p - 64u);
          bool good = t < T;
          for (int a = 0; good && a < kRegBits; ++a) good = o.t[a] != t;
          return good;
        };
        // a lane flip riding on the record (type word bits 8..): a 20-bit
        // word, bit 0 set, nothing above
        auto flip_ok = [&](unsigned f) {
          if (f == 0) return true;
          if (!(f & 1u) || (f >> 20) != 0 || !selector_ok((f >> 3) & 127u)) return false;
          if (f & 2048u) return selector_ok((f >> 12) & 127u);
          return ((f >> 12) & 255u) == 0;
        };
        if (sub[0] < 0 || sub[2] < 0 || !flip_ok((unsigned)sub[0] >> 8))
          return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: bad lane flip");
        const int ty = sub[0] & kRecType, cd = sub[1], cz = sub[2];
        // a selector-control word (SUB_CX): 1..3 selectors as for SUB_DIAGV,
        // their values in bits 23.., nothing above
        auto controls_ok = [&](unsigned w) {
          const unsigned m = w & 3u;
          bool good = m != 0 && (w >> (23 + m)) == 0 &&
                      ((w >> (2 + 7 * m)) & ((1u << (21 - 7 * m)) - 1u)) == 0;
          for (unsigned j = 0; good && j < m; ++j) good = selector_ok((w >> (2 + 7 * j)) & 127u);
          return good;
        };
        if (ty == SUB_DIAGV) {
          const unsigned c = (unsigned)cd, m = c & 3u, w = (unsigned)sub[3];
          // m selectors (none only with flip-mask bits, w >> 8)
          bool good = (m != 0 || ((w >> 8) & 15u)) && (c >> (2 + 7 * m)) == 0;
          for (unsigned j = 0; good && j < m; ++j) {
            const unsigned p = (c >> (2 + 7 * j)) & 127u;
            // a lane-group selector must be a tile bit outside the registers
            if (p >= 64u) {
              const int t = (int)(p - 64u);
              good = t < T;
              for (int a = 0; good && a < kRegBits; ++a) good = o.t[a] != t;
            }
          }
          // tables: 16 phases each, or compact (mask 16 | d): 2^popcount(d)
          if ((w & ~0xF1Fu) != 0 || (!(w & 16u) && (w & 15u))) good = false;
          const int per = (w & 16u) ? (1 << __builtin_popcount(w & 15u)) : kRegs;
          if (!good) return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: bad selector diagonal");
          if (cz + (per << m) > ncoef)
            return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: sub-op coefficients out of range");
          continue;
        }
        if (ty < SUB_D1 || ty > SUB_D1X || ncode[ty] == 0 || cd < 0 || cd >= ncode[ty])
          return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: bad sub-op");
        if (ty == SUB_CX) {
          const int cm = (cd >> 2) & 15, cv = (cd >> 6) & 15;
          if (((cm >> (cd & 3)) & 1) || (cv & ~cm))
            return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: bad CX control bits");
          // selector controls: none, or a selector-control word
          const unsigned w = (unsigned)sub[3];
          if (w != 0 && !controls_ok(w))
            return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: bad CX selector controls");
        }
        int nc = need[ty] << ((ty != SUB_CX && (cd & kD1Flip)) ? 1 : 0);
        if (ty == SUB_D1 || ty == SUB_D1R || ty == SUB_D1X) {
          // selectors: none, or m = w & 3 codes as SUB_DIAGV's, one set of
          // coefficients per selector value
          const unsigned w = (unsigned)sub[3], m = w & 3u;
          bool good = (w >> (2 + 7 * m)) == 0;
          for (unsigned j = 0; good && j < m; ++j) good = selector_ok((w >> (2 + 7 * j)) & 127u);
          if (!good) return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: bad 2x2 selectors");
          nc <<= m;
        }
        if (nc && cz + nc > ncoef)
          return qfx::set_error(QFX_EINVAL, "qfx_tile_check_ops: sub-op coefficients out of range");
      }
      continue;
    }
    if ((o.kind != OP_DENSE && o.kind != OP_SPARSE) || o.k < 1 || o.k > kmax || o.k > T)
      return qfx::set_error(QFX_EINVAL,
                            "qfx_tile_check_ops: bad op kind, or 4-5 qubit op without QFX_TILE_WIDE");
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
                  const int* dest, const void* ops, int nops, const void* coef, int ncoef,
                  const void* srcs, int block, int flags) {
  int rc = check_pass(n, tile_bits, tile_pos, nops, block, "qfx_tile_pass");
  if (rc) return rc;
  if (dest) {
    // in place: tile bits permute among the tile positions, the others stay
    uint64_t tmask = 0, seen = 0;
    for (int j = 0; j < tile_bits; ++j) tmask |= (uint64_t)1 << tile_pos[j];
    for (int b = 0; b < n; ++b) {
      const bool tb = (tmask >> b) & 1u;
      if (dest[b] < 0 || dest[b] >= n || ((seen >> dest[b]) & 1u) ||
          (tb ? !((tmask >> dest[b]) & 1u) : dest[b] != b))
        return qfx::set_error(QFX_EINVAL,
                              "qfx_tile_pass: dest must permute the tile positions among "
                              "themselves and keep every other index bit (the pass is in place)");
      seen |= (uint64_t)1 << dest[b];
    }
  }
  const bool wide = (flags & QFX_TILE_WIDE) != 0;
  const bool pers = (flags & QFX_TILE_PERSISTENT) != 0;
  const bool mf = (flags & QFX_TILE_MFMA) != 0;
  const bool reg_io = (flags & QFX_TILE_REG_IO) != 0;
  const bool reg_in = (flags & QFX_TILE_REG_IN) != 0;
  if (flags & ~(QFX_TILE_WIDE | QFX_TILE_PERSISTENT | QFX_TILE_MFMA | QFX_TILE_REG_IO |
                QFX_TILE_REG_IN))
    return qfx::set_error(QFX_EINVAL, "qfx_tile_pass: unknown flags");
  if (reg_in && (wide || mf || block > kRphaseMaxBlock || nops < 1 ||
                 (1 << tile_bits) != block * kRegs))
    return qfx::set_error(QFX_EINVAL,
                          "qfx_tile_pass: QFX_TILE_REG_IN needs 16 amplitudes per lane, "
                          "block <= 512, no wide ops or matrix-core blocks");
  if (reg_io && (wide || mf || block > kRphaseMaxBlock || nops < 1 ||
                 (1 << tile_bits) != block * kRegs))
    return qfx::set_error(QFX_EINVAL,
                          "qfx_tile_pass: QFX_TILE_REG_IO needs 16 amplitudes per lane, "
                          "block <= 512, no wide ops or matrix-core blocks");
  if (wide && block > kWideOpMaxBlock)
    return qfx::set_error(QFX_EINVAL, "qfx_tile_pass: 4-5 qubit ops need block <= 256");
  if (mf && block > kMfmaMaxBlock)
    return qfx::set_error(QFX_EINVAL, "qfx_tile_pass: matrix-core blocks need block <= 512");
  if (!state || (nops > 0 && (!ops || !coef)) || (ncoef > 0 && !coef))
    return qfx::set_error(QFX_EINVAL, "qfx_tile_pass: null pointer");
  if (ncoef < 0 || (!wide && !mf && ncoef > kMaxStaged))
    return qfx::set_error(QFX_ERANGE,
                          "qfx_tile_pass: more than 768 coefficients in a pass without wide ops");
  TilePos pos;
  std::memset(&pos, 0, sizeof(pos));
  for (int j = 0; j < tile_bits; ++j) {
    pos.s[j] = tile_pos[j];
    pos.d[j] = dest ? dest[tile_pos[j]] : tile_pos[j];
    pos.permuted |= pos.d[j] != pos.s[j];
  }
  hipStream_t s = qfx::as_stream(stream);
  auto* st = static_cast<v2d*>(state);
  auto* o = static_cast<const qfx_tile_op*>(ops);
  auto* c = static_cast<const v2d*>(coef);
  auto* sr = static_cast<const int*>(srcs);
#define QFX_PASS(BS, LB, W, M) \
  launch_pass<BS, LB, W, M>(s, st, n, tile_bits, pos, o, nops, c, ncoef, sr, pers, reg_io, reg_in)
  switch (block) {
    case 64:
      if (mf) return wide ? QFX_PASS(64, 6, true, true) : QFX_PASS(64, 6, false, true);
      return wide ? QFX_PASS(64, 6, true, false) : QFX_PASS(64, 6, false, false);
    case 256:
      if (mf) return wide ? QFX_PASS(256, 8, true, true) : QFX_PASS(256, 8, false, true);
      return wide ? QFX_PASS(256, 8, true, false) : QFX_PASS(256, 8, false, false);
    case 512: return mf ? QFX_PASS(512, 9, false, true) : QFX_PASS(512, 9, false, false);
    default: return QFX_PASS(1024, 10, false, false);
  }
#undef QFX_PASS
}

}  // extern "C"

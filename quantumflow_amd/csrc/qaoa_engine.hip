// Batched QAOA MaxCut engine: expectation values and adjoint gradients of
// many instances (graphs of different sizes) in one persistent launch.
//
// Replaces, for a whole batch at once, the reference's per-instance chain
//   qubo_circuit (quantumflow/qaoa.py:22-64) -> Circuit.run
//   (quantumflow/circuits.py:87-99) -> State.expectation(graph_cuts)
//   (quantumflow/states.py:131-147, quantumflow/qaoa.py:68-81)
// and the TensorFlow autograd gradient of examples/qaoa_maxcut.py:53-64.
//
// Per instance the state is |+>^n, then per layer l the cost layer
// prod_e ZZ(-w gamma_l / pi) = exp(i gamma_l (W/2 - C(z))) (W = sum of edge
// weights, C(z) the cut value) and the mixer RX(beta_l) on every qubit.
//
// Execution: every instance has a host-built program of passes.  A pass
// names a tile bit set S (|S| = T <= 13) and a list of stages; each of its
// 2^(n-T) tiles is one work item.  Global pass k runs the k-th pass of every
// instance; a grid barrier separates k from k+1.  Instances with n <= T run
// entirely in LDS inside one item.  Within a tile the mixer is applied as
// register butterflies over up to 4 bits at a time, the cost layer as a
// table lookup of exp(i gamma (W/2 - c)) by integer cut count c.
//
// Adjoint backward (mode QFX_QAOA_BACKWARD): starting from the stored final
// state psi and lambda = C psi, the stages are undone in reverse order; the
// gradient of <C> is
//   d/d beta_l  = Im sum_j <lambda|X_j|psi>          (any point of mixer l)
//   d/d gamma_l = 2 Im <lambda|C|psi> - W Im <lambda|psi>   (after cost l)
// with lambda, psi carried back by RX(-beta) and the conjugate phases.
#include <cmath>
#include <cstring>

#include "qfx_common.h"

namespace {

enum Stage : int {
  ST_INIT = 1,         // tile <- 2^(-n/2)
  ST_COST = 2,         // psi *= exp(i g (W/2 - C))
  ST_RX = 3,           // RX(beta) on the tile bits in mask
  ST_EXPECT = 4,       // acc += sum |psi|^2 C
  ST_MAKE_LAMBDA = 5,  // lambda <- C psi
  ST_RX_ADJ = 6,       // grad_beta += Im sum <lam|X_j|psi>; RX(-beta) on psi, lam
  ST_COST_ADJ = 7,     // grad_gamma += 2 Im <lam|C|psi> - W Im <lam|psi>; conj phases
  ST_COST_RX = 8,      // ST_COST then ST_RX of the same layer (phases in one RX chunk)
  ST_RX_LOAD = 10,     // ST_RX applied while the tile is loaded (first stage)
  ST_RX_STORE = 11,    // ST_RX applied while the tile is stored (last stage)
};

enum PassFlags : int { PF_LOAD_PSI = 1, PF_LOAD_LAM = 2, PF_STORE_PSI = 4, PF_STORE_LAM = 8 };

constexpr int kMaxT = 13;
constexpr int kMaxPLayers = QFX_QAOA_MAX_LAYERS;
constexpr int kAccum = 1 + 2 * kMaxPLayers;  // value, d/dbeta, d/dgamma
constexpr int kMaxEdgesTab = 512;  // cut counts 0..kMaxEdgesTab-1 tabulated
// A grid barrier gives up after kBarrierSeconds of wall-clock time.  One
// phase of the largest program qfx_qaoa_check accepts (n <= 30: a 16 GiB
// state read and written once) streams in well under 0.1 s, so a timeout
// means workgroups that never arrive, not a slow phase.
constexpr int kBarrierSeconds = 10;
constexpr int kMaxDevices = 64;

__device__ __forceinline__ unsigned swz(unsigned e) {
  return e ^ (((e >> 3) ^ (e >> 6) ^ (e >> 9) ^ (e >> 12)) & 7u);
}

__device__ __forceinline__ v2d cmulc(const v2d a, const v2d b) {  // conj(a) * b
  v2d r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.x * b.y - a.y * b.x;
  return r;
}

struct Ctx {
  const qfx_qaoa_inst* inst;
  const qfx_qaoa_pass* ps;
  uint64_t base;      // tile base index (non-tile bits)
  int T;
  int n;
  int p;
  const double* beta;
  const double* gamma;
  const uint16_t* cnt;  // cut counts of this instance (or null)
  const double* cval;   // weighted cut values (or null)
  uint64_t mmask;       // mirror instances: all n index bits; else 0
  int mshift;           // mirror instances: n - 1
};

// Where the amplitude of full-state index z is kept: a mirror instance keeps
// only the indices with bit n-1 clear, and one with it set is the
// complementary index (MaxCut states are symmetric under flipping every
// bit); other instances keep z itself.
__device__ __forceinline__ uint64_t stored(const Ctx& c, uint64_t z) {
  return z ^ (c.mmask & (0ull - ((z >> c.mshift) & 1ull)));
}

__device__ __forceinline__ double cut_of(const Ctx& c, uint64_t z) {
  return c.cnt ? (double)c.cnt[z] : c.cval[z];
}

// Cost phase of index z (unit weights) from the table tab[c] =
// exp(i g (W/2 - c)) by its cut count c.
__device__ __forceinline__ v2d cost_phase(const Ctx& c, const v2d* tab, uint32_t z) {
  return tab[c.cnt[z]];
}

// Global indices of a thread's tile elements e = threadIdx.x + k BS: the
// deposit of e's bits at the tile positions splits into the thread's part
// (bits below log2 BS, fixed per item) and k's part, which steps by one
// carry-add per element (the other index bits set to one so the carry
// crosses them).  Calls f(e, z) for every element the thread owns.
template <int BS, typename F>
__device__ __forceinline__ void for_elems(const Ctx& c, F&& f) {
  constexpr int kLb = __builtin_ctz(BS);
  const unsigned tsize = 1u << c.T;
  if (threadIdx.x >= tsize) return;
  uint64_t zl = c.base, hm = 0, h0 = 0;
#pragma unroll
  for (int j = 0; j < kMaxT; ++j) {
    if (j >= c.T) continue;
    const uint64_t bit = (uint64_t)1 << c.ps->pos[j];
    if (j < kLb) zl |= ((threadIdx.x >> j) & 1u) ? bit : 0;
    else hm |= bit;
    if (j == kLb) h0 = bit;
  }
  const int nk = tsize > (unsigned)BS ? (int)(tsize / BS) : 1;
  uint64_t zh = 0;
#pragma unroll 4
  for (int k = 0; k < nk; ++k) {
    f(threadIdx.x + (unsigned)k * BS, stored(c, zl | zh));
    zh = ((zh | ~hm) + h0) & hm;
  }
}

// Elements a thread holds at once in a load or store with RX: bits 0..3 of
// its element index k.
constexpr int kIoBits = 4;
constexpr int kIoRegs = 1 << kIoBits;

// RX(theta) on the bits of km (bit j <-> bit j of the register index) of
// kIoRegs register amplitudes: rx_regs with a run-time bit set.
__device__ __forceinline__ void rx_kregs(v2d (&v)[kIoRegs], unsigned km, double c, double s) {
#pragma unroll
  for (int j = 0; j < kIoBits; ++j) {
    if (!((km >> j) & 1u)) continue;
#pragma unroll
    for (int r = 0; r < kIoRegs; ++r) {
      if (r & (1 << j)) continue;
      const v2d x = v[r], y = v[r | (1 << j)];
      v2d nx, ny;
      nx.x = fma(c, x.x, s * y.y);
      nx.y = fma(c, x.y, -s * y.x);
      ny.x = fma(c, y.x, s * x.y);
      ny.y = fma(c, y.y, -s * x.x);
      v[r] = nx;
      v[r | (1 << j)] = ny;
    }
  }
}

// Tile load (LOAD) or store of the state with RX(theta) on tile bits held in
// registers on the way.  Thread t owns elements e = t + k BS; it takes them
// kIoRegs consecutive k at a time, so bits 0..3 of k (tile bits log2 BS ..
// +3) pair elements it holds.  km: those bits, as bits of k.
template <int BS, bool LOAD>
__device__ __forceinline__ void io_tile_rx(const Ctx& c, v2d* L, v2d* g, unsigned km,
                                           double cs, double sn) {
  constexpr int kLb = __builtin_ctz(BS);
  const unsigned tsize = 1u << c.T;
  if (threadIdx.x >= tsize) return;
  uint64_t zl = c.base, hm = 0, h0 = 0;
#pragma unroll
  for (int j = 0; j < kMaxT; ++j) {
    if (j >= c.T) continue;
    const uint64_t bit = (uint64_t)1 << c.ps->pos[j];
    if (j < kLb) zl |= ((threadIdx.x >> j) & 1u) ? bit : 0;
    else hm |= bit;
    if (j == kLb) h0 = bit;
  }
  const int nk = tsize > (unsigned)BS ? (int)(tsize / BS) : 1;
  uint64_t zh = 0;
  // the loop that touches HBM steps zh, once per element
  for (int k0 = 0; k0 < nk; k0 += kIoRegs) {
    v2d v[kIoRegs];
#pragma unroll
    for (int i = 0; i < kIoRegs; ++i) {
      const unsigned e = threadIdx.x + (unsigned)(k0 + i) * BS;
      v[i] = v2d{0.0, 0.0};
      if (LOAD) {
        if (k0 + i < nk) v[i] = g[stored(c, zl | zh)];
        zh = ((zh | ~hm) + h0) & hm;
      } else if (k0 + i < nk) {
        v[i] = L[swz(e)];
      }
    }
    rx_kregs(v, km, cs, sn);
#pragma unroll
    for (int i = 0; i < kIoRegs; ++i) {
      const unsigned e = threadIdx.x + (unsigned)(k0 + i) * BS;
      if (LOAD) {
        if (k0 + i < nk) L[swz(e)] = v[i];
      } else {
        if (k0 + i < nk) g[stored(c, zl | zh)] = v[i];
        zh = ((zh | ~hm) + h0) & hm;
      }
    }
  }
}

// Block-wide sum; result valid in thread 0.  red must hold >= 16 doubles.
template <int BS>
__device__ __forceinline__ double block_total(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BS / 64; ++w) t += red[w];
  return t;
}

// RX(theta) butterflies on register bits 0..R-1 of v (register bit j <->
// tile bit b[j]); c = cos(theta/2), s = sin(theta/2).
// new_x = c x - i s y ; new_y = c y - i s x
template <int R>
__device__ __forceinline__ void rx_regs(v2d (&v)[16], double c, double s) {
#pragma unroll
  for (int j = 0; j < R; ++j) {
#pragma unroll
    for (int r = 0; r < (1 << R); ++r) {
      if (r & (1 << j)) continue;
      const v2d x = v[r], y = v[r | (1 << j)];
      v2d nx, ny;
      nx.x = fma(c, x.x, s * y.y);
      nx.y = fma(c, x.y, -s * y.x);
      ny.x = fma(c, y.x, s * x.y);
      ny.y = fma(c, y.y, -s * x.x);
      v[r] = nx;
      v[r | (1 << j)] = ny;
    }
  }
}

// Im of sum over register pairs across bit j of conj(l_x) p_y + conj(l_y) p_x
template <int R>
__device__ __forceinline__ double xgrad_regs(const v2d (&l)[16], const v2d (&pv)[16]) {
  double g = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
#pragma unroll
    for (int r = 0; r < (1 << R); ++r) {
      if (r & (1 << j)) continue;
      const int q = r | (1 << j);
      g += cmulc(l[r], pv[q]).y + cmulc(l[q], pv[r]).y;
    }
  }
  return g;
}

// Apply RX to the tile bits in `mask` (chunks of <= 4 bits in registers);
// with ADJ also to lambda (angle -beta) and accumulate the beta gradient.
// With COST the chunk first multiplies its elements by their cost phases
// (unit-weight instances: tab[c] = exp(i g (W/2 - c))).  Each thread then
// needs its elements' state indices, z = base | deposit(group bits) |
// deposit(register bits), the last part fixed per chunk.
template <int BS, int R, bool ADJ, bool COST>
__device__ __forceinline__ double rx_chunk(v2d* L, v2d* LL, int T, const int* bits, double c,
                                           double s, const Ctx& cx, const v2d* tab) {
  const int ngroups = 1 << (T - R);
  int sorted[4];
#pragma unroll
  for (int j = 0; j < R; ++j) sorted[j] = bits[j];
  // ascending insertion order
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = a + 1; b < R; ++b)
      if (sorted[b] < sorted[a]) {
        const int t = sorted[a];
        sorted[a] = sorted[b];
        sorted[b] = t;
      }
  unsigned soff[1 << R];
  uint32_t zoff[1 << R];
#pragma unroll
  for (int r = 0; r < (1 << R); ++r) {
    unsigned o = 0;
    uint32_t zo = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      o |= ((unsigned)(r >> j) & 1u) << bits[j];
      if (COST && ((r >> j) & 1)) zo |= 1u << cx.ps->pos[bits[j]];
    }
    soff[r] = swz(o);
    zoff[r] = zo;
  }
  double g = 0.0;
  for (int grp = threadIdx.x; grp < ngroups; grp += BS) {
    unsigned lb = (unsigned)grp;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int b = sorted[j];
      lb = ((lb >> b) << (b + 1)) | (lb & ((1u << b) - 1u));
    }
    uint32_t zb = 0;
    if (COST) {
      zb = (uint32_t)cx.base;
#pragma unroll
      for (int t = 0; t < kMaxT; ++t)
        if (t < T && ((lb >> t) & 1u)) zb |= 1u << cx.ps->pos[t];
    }
    const unsigned slb = swz(lb);
    v2d v[16];
#pragma unroll
    for (int r = 0; r < (1 << R); ++r) v[r] = L[slb ^ soff[r]];
    if (COST) {
#pragma unroll
      for (int r = 0; r < (1 << R); ++r) {
        v[r] = cmul(cost_phase(cx, tab, (uint32_t)stored(cx, zb | zoff[r])), v[r]);
      }
    }
    if (ADJ) {
      v2d l[16];
#pragma unroll
      for (int r = 0; r < (1 << R); ++r) l[r] = LL[slb ^ soff[r]];
      g += xgrad_regs<R>(l, v);
      rx_regs<R>(l, c, s);
#pragma unroll
      for (int r = 0; r < (1 << R); ++r) LL[slb ^ soff[r]] = l[r];
    }
    rx_regs<R>(v, c, s);
#pragma unroll
    for (int r = 0; r < (1 << R); ++r) L[slb ^ soff[r]] = v[r];
  }
  return g;
}

// RX(theta) on the tile bits of mask; returns the beta-gradient terms (ADJ).
// With COST, the first chunk applies the cost phases before its butterflies.
// That chunk holds the highest (up to) four bits, so that neighbouring lanes
// differ in low state-index bits and their cut-table reads share cache
// lines; RX gates on different bits commute, so the chunk order is free.
template <int BS, bool ADJ, bool COST>
__device__ __forceinline__ double rx_mask(v2d* L, v2d* LL, int T, unsigned mask, double theta,
                                          const Ctx& cx, const v2d* tab) {
  const double c = cos(0.5 * theta), s = sin(0.5 * theta);
  double g = 0.0;
  bool first = true;
  while (mask) {
    int bits[4];
    int r = 0;
    if (COST && first) {
      while (mask && r < 4) {  // any order: register bit j <-> bits[j]
        const int b = 31 - __builtin_clz(mask);
        mask &= ~(1u << b);
        bits[r++] = b;
      }
    } else {
      while (mask && r < 4) {
        const int b = __builtin_ctz(mask);
        mask &= mask - 1;
        bits[r++] = b;
      }
    }
    if (COST && first) {
      switch (r) {
        case 1: g += rx_chunk<BS, 1, ADJ, true>(L, LL, T, bits, c, s, cx, tab); break;
        case 2: g += rx_chunk<BS, 2, ADJ, true>(L, LL, T, bits, c, s, cx, tab); break;
        case 3: g += rx_chunk<BS, 3, ADJ, true>(L, LL, T, bits, c, s, cx, tab); break;
        default: g += rx_chunk<BS, 4, ADJ, true>(L, LL, T, bits, c, s, cx, tab); break;
      }
    } else {
      switch (r) {
        case 1: g += rx_chunk<BS, 1, ADJ, false>(L, LL, T, bits, c, s, cx, tab); break;
        case 2: g += rx_chunk<BS, 2, ADJ, false>(L, LL, T, bits, c, s, cx, tab); break;
        case 3: g += rx_chunk<BS, 3, ADJ, false>(L, LL, T, bits, c, s, cx, tab); break;
        default: g += rx_chunk<BS, 4, ADJ, false>(L, LL, T, bits, c, s, cx, tab); break;
      }
    }
    first = false;
    __syncthreads();
  }
  return g;
}

// exp(i g (W/2 - c)) for c = 0..ncut into tab (conj if CONJ)
template <int BS>
__device__ void phase_table(v2d* tab, int ncut, double g, double wsum, bool conj) {
  for (int c = threadIdx.x; c <= ncut; c += BS) {
    double sn, cs;
    sincos(g * (0.5 * wsum - (double)c), &sn, &cs);
    v2d t;
    t.x = cs;
    t.y = conj ? -sn : sn;
    tab[c] = t;
  }
  __syncthreads();
}

// BWD selects the kernel of a mode: the forward kernel holds no adjoint
// stages, the backward kernel no fused cost chunk (its programs keep the
// cost stages separate).
template <int BS, bool BWD>
__device__ void run_item(const qfx_qaoa_inst& in, const qfx_qaoa_pass& ps, uint32_t tile,
                         const double* __restrict__ angles, v2d* __restrict__ psi_pool,
                         v2d* __restrict__ lam_pool, const uint16_t* __restrict__ cnt_pool,
                         const double* __restrict__ val_pool, double* __restrict__ partial,
                         int width, v2d* lds, double* red, double* accum, v2d* tab) {
  Ctx c;
  c.inst = &in;
  c.ps = &ps;
  c.T = ps.T;
  c.n = in.n;
  c.p = in.p;
  c.beta = angles + in.angle_off;
  c.gamma = angles + in.angle_off + in.p;
  c.cnt = in.weighted ? nullptr : cnt_pool + in.cut_off;
  c.cval = in.weighted ? val_pool + in.cut_off : nullptr;
  c.mmask = in.mirror ? ((uint64_t)1 << in.n) - 1 : 0;
  c.mshift = in.mirror ? in.n - 1 : 0;
  uint64_t base = tile;
#pragma unroll
  for (int j = 0; j < kMaxT; ++j) {
    if (j < c.T) {
      const int b = ps.pos[j];
      base = ((base >> b) << (b + 1)) | (base & (((uint64_t)1 << b) - 1));
    }
  }
  c.base = base;
  const unsigned tsize = 1u << c.T;
  v2d* L = lds;
  v2d* LL = lds + tsize;
  v2d* psi = psi_pool + in.psi_off;
  v2d* lam = lam_pool ? lam_pool + in.lam_off : nullptr;
  if (threadIdx.x < kAccum) accum[threadIdx.x] = 0.0;
  // RX bits applied on the way in from / out to HBM (forward programs): the
  // bits a thread holds in registers (kIoBits above log2 BS) there, the
  // others as an LDS chunk right after the load or before the store.
  constexpr int kLb = __builtin_ctz(BS);
  constexpr unsigned kRegMask = (kIoRegs - 1u) << kLb;
  int s_begin = 0, s_end = ps.nstage;
  unsigned ld_mask = 0, st_mask = 0;
  int ld_layer = 0, st_layer = 0;
  if (!BWD) {
    if (s_end > 0 && ps.stage[0][0] == ST_RX_LOAD) {
      ld_layer = ps.stage[0][1];
      ld_mask = (unsigned)ps.stage[0][2];
      s_begin = 1;
    }
    if (s_end > s_begin && ps.stage[s_end - 1][0] == ST_RX_STORE) {
      st_layer = ps.stage[s_end - 1][1];
      st_mask = (unsigned)ps.stage[s_end - 1][2];
      s_end -= 1;
    }
  }
  if (ps.flags & PF_LOAD_PSI) {
    if (!BWD && (ld_mask & kRegMask)) {
      const double th = 0.5 * c.beta[ld_layer];
      io_tile_rx<BS, true>(c, L, psi, (ld_mask & kRegMask) >> kLb, cos(th), sin(th));
    } else {
      for_elems<BS>(c, [&](unsigned e, uint64_t z) { L[swz(e)] = psi[z]; });
    }
  }
  if (BWD && (ps.flags & PF_LOAD_LAM))
    for_elems<BS>(c, [&](unsigned e, uint64_t z) { LL[swz(e)] = lam[z]; });
  __syncthreads();
  if (!BWD && (ld_mask & ~kRegMask))
    rx_mask<BS, false, false>(L, LL, c.T, ld_mask & ~kRegMask, c.beta[ld_layer], c, tab);
  const bool unit = !in.weighted;
  // Weighted instances run a fused stage as its two parts (their phases
  // come from sincos per amplitude, not from the cut-count table).
  for (int si = s_begin; si < s_end; ++si) {
    const int stype = ps.stage[si][0], layer = ps.stage[si][1];
    const unsigned mask = (unsigned)ps.stage[si][2];
    const bool split = !BWD && !unit && stype == ST_COST_RX;
    for (int part = 0; part < (split ? 2 : 1); ++part) {
      int type = stype;
      if (split) type = part == 0 ? ST_COST : ST_RX;
      switch (type) {
        case ST_INIT: {
          const double a = ldexp(1.0, -(c.n / 2)) * ((c.n & 1) ? 0.70710678118654752440 : 1.0);
          for (unsigned e = threadIdx.x; e < tsize; e += BS) L[swz(e)] = v2d{a, 0.0};
          __syncthreads();
          break;
        }
        case ST_COST:
        case ST_COST_ADJ: {
          const double g = c.gamma[layer];
          const bool adj = BWD && type == ST_COST_ADJ;
          if (unit) phase_table<BS>(tab, in.nedges, g, in.wsum, adj);
          double acc = 0.0;
          for_elems<BS>(c, [&](unsigned e, uint64_t z) {
            v2d ph;
            double cz;
            if (unit) {
              const int k = c.cnt[z];
              cz = (double)k;
              ph = tab[k];
            } else {
              cz = c.cval[z];
              double sn, cs;
              sincos(g * (0.5 * in.wsum - cz), &sn, &cs);
              ph = v2d{cs, adj ? -sn : sn};
            }
            const unsigned sl = swz(e);
            const v2d x = L[sl];
            if (adj) {
              const v2d l = LL[sl];
              const v2d lp = cmulc(l, x);  // conj(l) psi
              acc += 2.0 * cz * lp.y - in.wsum * lp.y;
              LL[sl] = cmul(ph, l);
            }
            L[sl] = cmul(ph, x);
          });
          if (adj) {
            const double t = block_total<BS>(acc, red);
            if (threadIdx.x == 0) accum[1 + c.p + layer] += t;
          }
          __syncthreads();
          break;
        }
        case ST_RX: {
          rx_mask<BS, false, false>(L, LL, c.T, mask, c.beta[layer], c, tab);
          break;
        }
        case ST_COST_RX: {  // unit weights, forward kernel only
          if constexpr (!BWD) {
            phase_table<BS>(tab, in.nedges, c.gamma[layer], in.wsum, false);
            rx_mask<BS, false, true>(L, LL, c.T, mask, c.beta[layer], c, tab);
          }
          break;
        }
        case ST_RX_ADJ: {
          if constexpr (BWD) {
            const double gsum =
                rx_mask<BS, true, false>(L, LL, c.T, mask, -c.beta[layer], c, tab);
            const double t = block_total<BS>(gsum, red);
            if (threadIdx.x == 0) accum[1 + layer] += t;
            __syncthreads();
          }
          break;
        }
        case ST_EXPECT: {
          double acc = 0.0;
          for_elems<BS>(c, [&](unsigned e, uint64_t z) {
            const v2d x = L[swz(e)];
            acc += fma(x.x, x.x, x.y * x.y) * cut_of(c, z);
          });
          const double t = block_total<BS>(acc, red);
          if (threadIdx.x == 0) accum[0] += t;
          __syncthreads();
          break;
        }
        case ST_MAKE_LAMBDA: {
          if constexpr (BWD) {
            for_elems<BS>(c, [&](unsigned e, uint64_t z) {
              const unsigned sl = swz(e);
              const double cz = cut_of(c, z);
              const v2d x = L[sl];
              LL[sl] = v2d{cz * x.x, cz * x.y};
            });
            __syncthreads();
          }
          break;
        }
      }
    }  // part
  }
  if (!BWD && (st_mask & ~kRegMask))
    rx_mask<BS, false, false>(L, LL, c.T, st_mask & ~kRegMask, c.beta[st_layer], c, tab);
  if (ps.flags & PF_STORE_PSI) {
    if (!BWD && (st_mask & kRegMask)) {
      const double th = 0.5 * c.beta[st_layer];
      io_tile_rx<BS, false>(c, L, psi, (st_mask & kRegMask) >> kLb, cos(th), sin(th));
    } else {
      for_elems<BS>(c, [&](unsigned e, uint64_t z) { psi[z] = L[swz(e)]; });
    }
  }
  if (BWD && (ps.flags & PF_STORE_LAM))
    for_elems<BS>(c, [&](unsigned e, uint64_t z) { lam[z] = LL[swz(e)]; });
  __syncthreads();
  if ((int)threadIdx.x < width) partial[threadIdx.x] = accum[threadIdx.x];
  __syncthreads();
}

// Grid barrier: monotone counter, wait bounded by the device's constant-rate
// wall clock (limit_ticks); on timeout it sets this launch's error flag
// (err[0], which makes every waiting workgroup leave too) and the sticky flag
// (err[1], cleared only by the batch's owner), and returns false.
__device__ bool grid_barrier(unsigned* counter, unsigned target, int* err,
                             uint64_t limit_ticks) {
  __syncthreads();
  __shared__ int s_ok;
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(counter, 1u);
    const uint64_t t0 = wall_clock64();
    int ok = 1;
    while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(8);
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
          wall_clock64() - t0 > limit_ticks) {
        atomicExch(err, 1);
        atomicOr(err + 1, 1);
        ok = 0;
        break;
      }
    }
    __threadfence();
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

template <int BS, bool BWD>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2))) void qaoa_kernel(
    const qfx_qaoa_inst* __restrict__ insts, const qfx_qaoa_pass* __restrict__ passes,
    const qfx_qaoa_item* __restrict__ items, const int* __restrict__ kstart, int nk,
    const double* __restrict__ angles, v2d* __restrict__ psi_pool, v2d* __restrict__ lam_pool,
    const uint16_t* __restrict__ cnt_pool, const double* __restrict__ val_pool,
    double* __restrict__ partials, int width, unsigned* sync, int* err, uint64_t limit_ticks,
    uint64_t* __restrict__ timing, int64_t timing_len) {
  extern __shared__ v2d lds[];
  __shared__ double red[16];
  __shared__ double accum[kAccum];
  __shared__ v2d tab[kMaxEdgesTab];
  for (int k = 0; k < nk; ++k) {
    const int64_t tslot = 2 * ((int64_t)k * gridDim.x + blockIdx.x);
    const bool timed = timing != nullptr && tslot + 1 < timing_len && threadIdx.x == 0;
    if (timed) timing[tslot] = wall_clock64();
    for (int it = kstart[k] + (int)blockIdx.x; it < kstart[k + 1]; it += gridDim.x) {
      const qfx_qaoa_item item = items[it];
      const qfx_qaoa_inst& in = insts[item.inst];
      run_item<BS, BWD>(in, passes[item.pass], (uint32_t)item.tile, angles, psi_pool,
                        lam_pool, cnt_pool, val_pool,
                        partials + (uint64_t)item.partial * width, width, lds, red, accum, tab);
    }
    if (timed) timing[tslot + 1] = wall_clock64();
    if (k + 1 < nk && !grid_barrier(sync, (unsigned)(k + 1) * gridDim.x, err, limit_ticks))
      return;
  }
}

// out[i * width + j] = sum of the instance's partial slots: one workgroup
// per output, a fixed strided split over its threads and a fixed-order tree
// (bitwise reproducible).
constexpr int kReduceBlock = 256;

__global__ __launch_bounds__(kReduceBlock) void qaoa_reduce_kernel(
    const qfx_qaoa_inst* __restrict__ insts, int ninst, const double* __restrict__ partials,
    int width, double* __restrict__ out) {
  const int i = blockIdx.x;
  const int inst = i / width, j = i % width;
  const qfx_qaoa_inst& in = insts[inst];
  double s = 0.0;
  for (int q = threadIdx.x; q < in.nslots; q += kReduceBlock)
    s += partials[(uint64_t)(in.slot0 + q) * width + j];
  __shared__ double red[kReduceBlock / 64];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  // a mirror instance's items each stand for their tile and its complement
  if (threadIdx.x == 0) out[i] = ((red[0] + red[1]) + (red[2] + red[3])) * (in.mirror ? 2.0 : 1.0);
}

__global__ void zero_sync_kernel(unsigned* sync, int* err) {
  if (threadIdx.x == 0) {
    sync[0] = 0u;
    err[0] = 0;
  }
}

// Persistent grid of one instantiation on one device: every workgroup that
// fits the device at once (occupancy x CUs for this block size, mode and LDS
// size), so all of them are resident together and the grid barrier cannot
// wait on a workgroup that has not started.  Cached per device and LDS size.
struct GridCache {
  int grid = 0;
  size_t lds = 0;
  uint64_t limit_ticks = 0;
};

template <int BS, bool BWD>
int launch_qaoa(hipStream_t s, const qfx_qaoa_args& a, int T) {
  const size_t lds = (sizeof(v2d) << T) * (BWD ? 2 : 1);
  static GridCache cache[kMaxDevices];
  int dev = 0;
  int rc = qfx::check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return qfx::set_error(QFX_ERANGE, "qfx_qaoa_run: device index out of range");
  GridCache& gc = cache[dev];
  if (gc.grid == 0 || gc.lds != lds) {
    rc = qfx::check_hip(hipFuncSetAttribute((const void*)qaoa_kernel<BS, BWD>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                        "hipFuncSetAttribute");
    if (rc) return rc;
    int per_cu = 0, cus = 0, rate_khz = 0;
    rc = qfx::check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                            &per_cu, (const void*)qaoa_kernel<BS, BWD>, BS, lds),
                        "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (rc) return rc;
    if (per_cu < 1) return qfx::set_error(QFX_ERANGE, "qfx_qaoa_run: kernel does not fit a CU");
    rc = qfx::check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                        "hipDeviceGetAttribute");
    if (rc) return rc;
    rc = qfx::check_hip(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev),
                        "hipDeviceGetAttribute");
    if (rc) return rc;
    if (rate_khz <= 0) return qfx::set_error(QFX_EKERNEL, "qfx_qaoa_run: no wall-clock rate");
    gc.grid = per_cu * cus;
    gc.lds = lds;
    gc.limit_ticks = (uint64_t)rate_khz * 1000u * kBarrierSeconds;
  }
  const int grid = gc.grid;
  int g = grid;
  if (a.max_items_per_k > 0 && a.max_items_per_k < g) g = a.max_items_per_k;
  auto* sync = static_cast<unsigned*>(a.sync);
  auto* err = reinterpret_cast<int*>(sync + 1);
  hipLaunchKernelGGL(zero_sync_kernel, dim3(1), dim3(64), 0, s, sync, err);
  rc = qfx::check_launch("zero_sync_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL((qaoa_kernel<BS, BWD>), dim3(g), dim3(BS), lds, s,
                     static_cast<const qfx_qaoa_inst*>(a.insts),
                     static_cast<const qfx_qaoa_pass*>(a.passes),
                     static_cast<const qfx_qaoa_item*>(a.items), static_cast<const int*>(a.kstart),
                     a.nk, static_cast<const double*>(a.angles), static_cast<v2d*>(a.psi_pool),
                     static_cast<v2d*>(a.lam_pool), static_cast<const uint16_t*>(a.cnt_pool),
                     static_cast<const double*>(a.val_pool), static_cast<double*>(a.partials),
                     a.width, sync, err, gc.limit_ticks, static_cast<uint64_t*>(a.timing),
                     a.timing ? a.timing_len : (int64_t)0);
  rc = qfx::check_launch("qaoa_kernel");
  if (rc) return rc;
  const int nout = a.ninst * a.width;
  hipLaunchKernelGGL(qaoa_reduce_kernel, dim3(nout), dim3(kReduceBlock), 0, s,
                     static_cast<const qfx_qaoa_inst*>(a.insts), a.ninst,
                     static_cast<const double*>(a.partials), a.width,
                     static_cast<double*>(a.out));
  return qfx::check_launch("qaoa_reduce_kernel");
}

}  // namespace

extern "C" {

int qfx_qaoa_check(const qfx_qaoa_inst* insts, int ninst, const qfx_qaoa_pass* passes,
                   int npasses, const qfx_qaoa_item* items, const int* kstart, int nk, int width,
                   int64_t psi_pool_len, int64_t lam_pool_len, int64_t cnt_pool_len,
                   int64_t val_pool_len, int64_t npartial, int mode, int tile_bits) {
  const int tdef = mode == QFX_QAOA_BACKWARD ? 12 : 13;
  if (tile_bits < 0 || tile_bits > tdef)
    return qfx::set_error(QFX_ERANGE, "qfx_qaoa_check: tile_bits out of range");
  const int tmax = tile_bits ? tile_bits : tdef;
  for (int i = 0; i < ninst; ++i) {
    const qfx_qaoa_inst& in = insts[i];
    if (in.n < 1 || in.n > 30 || in.p < 1 || in.p > kMaxPLayers)
      return qfx::set_error(QFX_ERANGE, "qfx_qaoa_check: n or p out of range");
    const int64_t len = (int64_t)1 << in.n;
    if (in.weighted) {
      if (in.cut_off < 0 || in.cut_off + len > val_pool_len)
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: cut values out of range");
    } else {
      if (in.cut_off < 0 || in.cut_off + len > cnt_pool_len)
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: cut counts out of range");
      if (in.nedges < 0 || in.nedges >= kMaxEdgesTab)
        return qfx::set_error(QFX_ERANGE, "qfx_qaoa_check: too many edges for the phase table");
    }
    if (in.mirror != 0 && in.mirror != 1)
      return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: mirror must be 0 or 1");
    const int64_t slen = in.mirror ? len / 2 : len;   // amplitudes kept
    if (in.psi_off < 0 || in.psi_off + slen > psi_pool_len)
      return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: state out of range");
    if (mode == QFX_QAOA_BACKWARD && (in.lam_off < 0 || in.lam_off + slen > lam_pool_len))
      return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: lambda out of range");
    if (in.slot0 < 0 || in.nslots < 0 || in.slot0 + in.nslots > npartial)
      return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: partial slots out of range");
  }
  if (width < 1 || width > kAccum)
    return qfx::set_error(QFX_ERANGE, "qfx_qaoa_check: width out of range");
  for (int q = 0; q < npasses; ++q) {
    const qfx_qaoa_pass& ps = passes[q];
    if (ps.T < 1 || ps.T > tmax || ps.nstage < 0 || ps.nstage > QFX_QAOA_MAX_STAGES)
      return qfx::set_error(QFX_ERANGE, "qfx_qaoa_check: pass tile or stage count out of range");
    if (mode != QFX_QAOA_BACKWARD && (ps.flags & (PF_LOAD_LAM | PF_STORE_LAM)))
      return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: lambda traffic in a forward program");
    for (int j = 0; j < ps.T; ++j)
      if (ps.pos[j] < 0 || ps.pos[j] > 30 || (j && ps.pos[j] <= ps.pos[j - 1]))
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: tile positions not ascending");
    for (int si = 0; si < ps.nstage; ++si) {
      const int ty = ps.stage[si][0], l = ps.stage[si][1];
      const unsigned mask = (unsigned)ps.stage[si][2];
      if (ty < ST_INIT || (ty > ST_COST_RX && ty != ST_RX_LOAD && ty != ST_RX_STORE) || l < 0 ||
          l >= kMaxPLayers)
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: bad stage");
      const bool io = ty == ST_RX_LOAD || ty == ST_RX_STORE;
      const bool rx = ty == ST_RX || ty == ST_RX_ADJ || ty == ST_COST_RX || io;
      if (io && (mode == QFX_QAOA_BACKWARD || mask == 0))
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: bad load/store RX stage");
      if (ty == ST_RX_LOAD && (si != 0 || !(ps.flags & PF_LOAD_PSI)))
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: load RX stage not first or no load");
      if (ty == ST_RX_STORE && (si != ps.nstage - 1 || !(ps.flags & PF_STORE_PSI)))
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: store RX stage not last or no store");
      if (rx && (mask >> ps.T))
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: RX mask outside the tile");
      if (ty == ST_COST_RX && mask == 0)
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: fused cost stage without RX bits");
      if (ty == ST_COST_RX && mode == QFX_QAOA_BACKWARD)
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: fused cost stage in a backward program");
      if (mode != QFX_QAOA_BACKWARD &&
          (ty == ST_MAKE_LAMBDA || ty == ST_RX_ADJ || ty == ST_COST_ADJ))
        return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: backward stage in a forward program");
    }
  }
  if (nk < 0 || kstart[0] != 0)
    return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: bad pass index table");
  for (int k = 0; k < nk; ++k)
    if (kstart[k + 1] < kstart[k]) return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: kstart");
  for (int it = 0; it < kstart[nk]; ++it) {
    const qfx_qaoa_item& item = items[it];
    if (item.inst < 0 || item.inst >= ninst || item.pass < 0 || item.pass >= npasses ||
        item.partial < 0 || item.partial >= npartial)
      return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: item out of range");
    const qfx_qaoa_inst& in = insts[item.inst];
    const qfx_qaoa_pass& ps = passes[item.pass];
    if (ps.T > in.n || item.tile < 0 || ((int64_t)item.tile >> (in.n - ps.T)) != 0)
      return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: tile out of range");
    // mirror: bit n-1 is the pass's highest tile bit, the tile's highest
    // non-tile bit is clear (so its complement tile is not an item)
    if (in.mirror && (ps.T < 1 || ps.pos[ps.T - 1] != in.n - 1 || in.n - ps.T < 1 ||
                      ((int64_t)item.tile >> (in.n - ps.T - 1)) != 0))
      return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: mirror pass or tile out of range");
    for (int j = 0; j < ps.T; ++j)
      if (ps.pos[j] >= in.n) return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: tile bit >= n");
    for (int si = 0; si < ps.nstage; ++si)
      if (ps.stage[si][1] >= in.p) return qfx::set_error(QFX_EINVAL, "qfx_qaoa_check: layer >= p");
  }
  return QFX_OK;
}

int qfx_qaoa_run(void* stream, const qfx_qaoa_args* args) {
  if (!args) return qfx::set_error(QFX_EINVAL, "qfx_qaoa_run: null args");
  const qfx_qaoa_args& a = *args;
  if (!a.insts || !a.passes || !a.items || !a.kstart || !a.angles || !a.psi_pool ||
      !a.partials || !a.out || !a.sync || (!a.cnt_pool && !a.val_pool))
    return qfx::set_error(QFX_EINVAL, "qfx_qaoa_run: null pointer");
  if (a.mode == QFX_QAOA_BACKWARD && !a.lam_pool)
    return qfx::set_error(QFX_EINVAL, "qfx_qaoa_run: backward needs lam_pool");
  if (a.width < 1 || a.width > kAccum || a.ninst < 1 || a.nk < 1)
    return qfx::set_error(QFX_ERANGE, "qfx_qaoa_run: sizes out of range");
  const int tdef = a.mode == QFX_QAOA_BACKWARD ? 12 : 13;
  if (a.tile_bits < 0 || a.tile_bits > tdef)
    return qfx::set_error(QFX_ERANGE, "qfx_qaoa_run: tile_bits out of range");
  const int T = a.tile_bits ? a.tile_bits : tdef;
  hipStream_t s = qfx::as_stream(stream);
  const bool bwd = a.mode == QFX_QAOA_BACKWARD;
  switch (a.block) {
    case 256: return bwd ? launch_qaoa<256, true>(s, a, T) : launch_qaoa<256, false>(s, a, T);
    case 512: return bwd ? launch_qaoa<512, true>(s, a, T) : launch_qaoa<512, false>(s, a, T);
    case 1024: return bwd ? launch_qaoa<1024, true>(s, a, T) : launch_qaoa<1024, false>(s, a, T);
  }
  return qfx::set_error(QFX_EINVAL, "qfx_qaoa_run: block must be 256, 512 or 1024");
}

int qfx_qaoa_status(const void* sync, int* failed) {
  if (!sync || !failed) return qfx::set_error(QFX_EINVAL, "qfx_qaoa_status: null pointer");
  int e[2] = {0, 0};
  int rc = qfx::check_hip(hipMemcpy(e, static_cast<const unsigned*>(sync) + 1, sizeof(e),
                                    hipMemcpyDeviceToHost),
                          "hipMemcpy");
  if (rc) return rc;
  *failed = (e[0] | e[1]) != 0;
  return QFX_OK;
}

}  // extern "C"

/*
 * qfx.h -- C-ABI of the MI355X (gfx950) state-vector engine behind
 * quantumflow_amd.backend (the "rocm" QuantumFlow backend).
 *
 * Conventions
 *   - A state of n qubits is 2^n complex128 amplitudes, interleaved
 *     (re, im) doubles, contiguous, in device memory.  Flat index bit
 *     (n-1-axis) belongs to tensor axis `axis` (big-endian in axis order,
 *     quantumflow/qubits.py:63-68).
 *   - `targets[j]` is the index BIT POSITION of gate qubit j, i.e.
 *     n-1-axis_j.  Gate qubit 0 is the most significant bit of the gate's
 *     row/column index (quantumflow/qubits.py:70-80).
 *   - Matrices are row-major 2^k x 2^k complex128, row = output.
 *   - Host pointers are marked (host); everything else is device memory
 *     owned by the caller.  The library never frees caller memory.
 *   - `stream` is a hipStream_t (NULL = default stream).  Calls only
 *     enqueue work; none synchronises.
 *   - Every function returns 0 on success or a negative QFX_E* code and
 *     never throws; qfx_last_error() gives the message (thread-local).
 *   - out == in is allowed for every apply call (each work item reads its
 *     whole amplitude group before writing it).
 */
#ifndef QFX_H
#define QFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QFX_OK 0
#define QFX_EINVAL (-1)   /* bad argument (maps to ValueError)       */
#define QFX_ERANGE (-2)   /* size out of supported range             */
#define QFX_EHIP (-3)     /* HIP runtime error                       */
#define QFX_ENODEV (-4)   /* no usable gfx950 device                 */
#define QFX_EKERNEL (-5)  /* a kernel reported failure via its flag  */

#define QFX_MAX_QUBITS 34 /* 2^34 amplitudes = 256 GiB, fits 288 GB HBM */
#define QFX_MAX_DENSE_K 12

typedef struct {
  double re, im;
} qfx_c128;

/* ---- library ---------------------------------------------------------- */

/* Text of the last error on this thread (never NULL). */
const char* qfx_last_error(void);

/* Library version string. */
const char* qfx_version(void);

/* Initialise for the current device; returns QFX_ENODEV if it is not gfx950.
 * Safe to call more than once.  Fills (host) props[8] when props != NULL:
 * {CUs, LDS bytes per block, warp size, L2 bytes, clock kHz, total MiB,
 *  max threads per block, device wall-clock (s_memrealtime) rate kHz}. */
int qfx_init(int64_t* props);

/* ---- gate application (replaces numpybk.tensormul,
 *      quantumflow/backend/numpybk.py:159-214, on the Gate.run path
 *      quantumflow/ops.py:161-166) ---------------------------------------- */

/* Dense k-qubit gate, 1 <= k <= 3, matrix passed by value from the host. */
int qfx_apply_matrix(void* stream, void* out, const void* in, int n, int k,
                     const int* targets /*host,k*/,
                     const qfx_c128* matrix /*host,4^k*/);

/* Dense k-qubit gate, 1 <= k <= QFX_MAX_DENSE_K, matrix in device memory. */
int qfx_apply_matrix_dev(void* stream, void* out, const void* in, int n,
                         int k, const int* targets /*host,k*/,
                         const void* matrix /*device,4^k complex128*/);

/* Diagonal k-qubit gate (RZ, PHASE, CZ, CPHASE, ZZ, ...), 1 <= k <= 5:
 * out[i] = diag[r(i)] * in[i]. */
int qfx_apply_diagonal(void* stream, void* out, const void* in, int n, int k,
                       const int* targets /*host,k*/,
                       const qfx_c128* diag /*host,2^k*/);

/* Monomial k-qubit gate (X, CNOT, SWAP, ISWAP, CCNOT, CSWAP, ...),
 * 1 <= k <= 5: row r of the gate has its single non-zero phase[r] in
 * column src[r].  phase == NULL means all ones: the result is then a pure
 * permutation of the input, bit for bit. */
int qfx_apply_monomial(void* stream, void* out, const void* in, int n, int k,
                       const int* targets /*host,k*/,
                       const int* src /*host,2^k*/,
                       const qfx_c128* phase /*host,2^k or NULL*/);

/* In-place variants that only touch the rows a diagonal or monomial gate
 * changes (identity rows are skipped).  state is read and written. */
int qfx_apply_diagonal_inplace(void* stream, void* state, int n, int k,
                               const int* targets, const qfx_c128* diag);
int qfx_apply_monomial_inplace(void* stream, void* state, int n, int k,
                               const int* targets, const int* src,
                               const qfx_c128* phase);

/* ---- readouts (quantumflow/states.py:113-147, numpybk.py:125-128,
 *      quantumflow/qubits.py:180-182) ------------------------------------- */

/* Bytes of device workspace the reductions below need for `len` elements. */
int64_t qfx_reduce_workspace_bytes(int64_t len);

/* result[0] = sum_i conj(a_i) b_i   (np.vdot), complex128 in device memory */
int qfx_inner(void* stream, void* result, const void* a, const void* b,
              int64_t len, void* workspace);
/* out[i] = |in_i| (hypot), float64 */
int qfx_abs(void* stream, void* out, const void* in, int64_t len);
/* out[i] = |in_i|^2, float64 */
int qfx_abs2(void* stream, void* out, const void* in, int64_t len);
/* result[0] = sum_i |psi_i|^2 * diag_i   (diag float64, device) */
int qfx_expect_diag(void* stream, void* result, const void* psi,
                    const void* diag, int64_t len, void* workspace);
/* out[a] = <Z_a> = sum_i |psi_i|^2 (1 - 2 bit_{n-1-a}(i)) for every axis a */
int qfx_expect_z(void* stream, void* out, const void* psi, int n,
                 void* workspace);
/* out[0] = sum |psi_i|^2 over i with bit n-1-axis clear (qubit axis in |0>),
 * out[1] = the same over i with it set; float64.  out[0] is
 * P0(q).run(ket).norm() of Measure.run (quantumflow/stdops.py:53-54). */
int qfx_qubit_probs(void* stream, void* out, const void* psi, int n, int axis,
                    void* workspace);
/* result[0] = sum_i |psi_i|^2 (float64) */
int qfx_norm2(void* stream, void* result, const void* psi, int64_t len,
              void* workspace);

/* qfx_expect_diag without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Its block partials live in a 16 KiB array of the library's
 * own device memory (one copy per GPU, held until the process ends), so it
 * calls no allocator; len <= 2^34.  Calls that overlap on one GPU share
 * that array and race: order them on one stream. */
int qfx_expect_diag_noalloc(void* stream, void* result, const void* psi,
                            const void* diag, int64_t len);

/* ---- states ------------------------------------------------------------ */

/* Qubit relabelling (QubitVector.permute, quantumflow/qubits.py:145-161):
 * out[i] = in[j] where bit src_bit[b] of j equals bit b of i.  out != in. */
int qfx_permute_qubits(void* stream, void* out, const void* in, int n,
                       const int* src_bit /*host,n*/);

/* psi = |0...0> on n qubits. */
int qfx_zero_state(void* stream, void* psi, int n);
/* psi = |+...+> = 2^(-n/2) everywhere. */
int qfx_plus_state(void* stream, void* psi, int n);

/* ---- sampling the computational basis -----------------------------------
 * Replaces the numpy draws of State.sample (quantumflow/states.py:121-129,
 * np.random.multinomial over probabilities()) and State.measure
 * (quantumflow/states.py:149-160, np.random.choice): out[t] is a basis index
 * drawn with probability |psi_i|^2 / sum_j |psi_j|^2, from a SplitMix64
 * stream keyed by (seed, t); reproducible for equal (psi, seed).  Indices of
 * zero probability are never drawn.  Index bit n-1-a is qubit axis a. */
/* Bytes of device workspace qfx_sample needs for n qubits (0 if n invalid). */
int64_t qfx_sample_workspace_bytes(int n);
int qfx_sample(void* stream, int64_t* out, const void* psi, int n,
               int64_t trials, uint64_t seed, void* workspace);
/* counts[idx[i]] += 1 for i < m (counts: 2^n int64, zeroed by the caller;
 * every idx[i] must lie in [0, 2^n)). */
int qfx_histogram(void* stream, int64_t* counts, const int64_t* idx, int64_t m,
                  int n);

/* ---- density matrices (quantumflow_amd/channels.py) --------------------
 * rho: device complex128 [2^n x 2^n], row-major (row = ket index), the
 * vectorised 2n-qubit tensor of the reference's Density
 * (quantumflow/states.py:264-344).  Gates, Kraus operators and channels act
 * on it through qfx_apply_* (rows by U, columns by conj(U)); these add the
 * rest.  0 <= n <= QFX_DENSITY_MAX_QUBITS (1 <= n for the partial trace). */
#define QFX_DENSITY_MAX_QUBITS 15
/* rho = psi psi^H; replaces State.asdensity (quantumflow/states.py:162-165) */
int qfx_density_outer(void* stream, void* rho, const void* psi, int n);
/* rho = scale I; replaces mixed_density (quantumflow/states.py:322-326) */
int qfx_density_identity(void* stream, void* rho, int n, double scale);
/* out[i] = Re rho_ii (float64 [2^n]); replaces Density.probabilities
 * (quantumflow/states.py:302-305) */
int qfx_density_diag(void* stream, void* out, const void* rho, int n);
/* out = sum_i rho_ii (one complex128, deterministic); replaces
 * Density.trace (quantumflow/states.py:277-279) */
int qfx_density_trace(void* stream, void* out, const void* rho, int n);
/* y += w x over len complex128 elements; the weighted sum of
 * Kraus.evolve (quantumflow/channels.py:74-80) */
int qfx_density_axpy(void* stream, void* y, const void* x, double w, int64_t len);
/* out (device complex128 [2^m x 2^m], m = popcount(keep_mask)) = the
 * partial trace of rho over the index bits outside keep_mask (bit b <->
 * axis n-1-b), kept axes in their order; replaces QubitVector.partial_trace
 * (quantumflow/qubits.py:201-229).  keep_mask != 0, keep_mask < 2^n. */
int qfx_density_partial_trace(void* stream, void* out, const void* rho, int n,
                              uint32_t keep_mask);

/* ---- fused circuit execution: LDS-tiled passes --------------------------
 * Replaces the per-gate loop of Circuit.run (quantumflow/circuits.py:87-99)
 * for runs of gates that act inside one tile bit set.  A pass loads every
 * tile of 2^tile_bits amplitudes (index bits tile_pos[0..T-1], ascending,
 * which should include the lowest index bits for coalescing) into LDS,
 * applies `nops` ops to it and stores it back, in place. */

typedef struct {
  int32_t kind;       /* 0 dense (coef: 4^k row-major), 1 sparse (coef: 2^k
                         phases; srcs: 2^k source rows), 2 register phase,
                         3 dense block on the matrix cores (coef: 2*4^k
                         slots holding the packed real form, see below) */
  int32_t k;          /* dense/sparse: gate width, 1..5; matrix-core block:
                         3..5 */
  int32_t coef;       /* dense/sparse: first coefficient (complex128 index);
                         register phase: number of sub-ops */
  int32_t src;        /* sparse: first source-row index into srcs;
                         register phase: first sub-op (4 ints each) in srcs */
  int32_t t[5];       /* dense/sparse: tile bit of gate qubit j (MSB first);
                         register phase: tile bit of register bit j, j < 4 */
  int32_t sorted[5];  /* t[0..k-1] (register phase: t[0..3]) ascending */
  uint32_t active;    /* sparse: rows written */
  uint32_t phase;     /* sparse: rows multiplied by their phase */
  uint32_t lds[10];   /* register phase: LDS slot offsets derived from t and
                         the two maps (filled by qfx_tile_prepare_ops; other
                         kinds: zero) */
} qfx_tile_op;

/* A register phase keeps, per lane, the 16 amplitudes spanned by its four
 * register bits in registers and applies its sub-ops there.  Sub-op records
 * are 4 ints {type, position code, coef index, mask}, starting at a multiple
 * of 4 ints in srcs (srcs holds at least one record at `src`, also for a
 * phase without sub-ops: the first record is read ahead):
 *   1  2x2 matrix (4 coefs) on register bit `code`; with selectors in
 *      `mask` (m = mask & 3 codes as type 9's) one matrix per selector
 *      value v, at coef index 4 v (types 7 and 8 alike);
 *   7  real 2x2 matrix (real parts of the 4 coefs are read);
 *   8  2x2 matrix with real diagonal and imaginary off-diagonal (real parts
 *      of coefs 0 and 3, imaginary parts of coefs 1 and 2 are read);
 *   4  16 phases, one per register slot (exactly 1 where `mask` has no
 *      bit);
 *   5  X on register bit code & 3 where the register bits of the mask
 *      (code >> 2) & 15 equal the values (code >> 6) & 15 (X, CNOT, CCNOT
 *      and their negated-control forms as register exchanges); `mask` may
 *      add selector controls: m = mask & 3 selectors coded as type 9's,
 *      selector j required to equal bit 23 + j of `mask`;
 *   9  2^m tables of 16 phases (table v at coef index 16 v), m = code & 3
 *      in 0..3, applied as type 4 with table v, where bit j of v is selector
 *      p = (code >> (2 + 7 j)) & 127: the tile's state-index bit p (p < 64,
 *      a bit outside the tile) or the lane group's tile bit p - 64 (a tile
 *      bit outside the phase's registers).  Diagonal gates need none of
 *      their qubits in registers or in the tile.  `mask` = 16 | d: compact
 *      tables of 2^popcount(d) phases, slot r reading the entry of its bits
 *      under d; bits 8..11 of `mask`: bits f of the lane's flip mask (below)
 *      the table reads, slot r taking the entry of slot r ^ (flips & f)
 *      (m = 0 only with such bits).
 * Lane flips: any record may carry a controlled X gate whose controls are
 * all selectors, applied before it without moving amplitudes: a 20-bit
 * word in bits 8.. of the type (bit 0 set; bits 1-2 the register bit; bits 3-9 a selector coded as type
 * 9's and bit 10 its required value; bit 11 a second selector in bits
 * 12-18, value in bit 19).  A lane whose selectors hold records the X in its
 * flip mask f: register r then holds what register r ^ f holds without
 * flips, and the phase stores it there.  Records after a flip that read a
 * flipped bit carry a flag: types 1/7/8 code | 4 with 8 coefs per matrix
 * (m, then X m X, read on lanes where the bit is flipped; matrix v at 8 v), type 5 code | 1 << 10
 * (control values compared through f); diagonals are type 9 records that
 * name the flip-mask bits they read.
 * A register phase's `phase` and `active` words are its load and store
 * maps: register slot r holds logical phase index c ^ (XOR of column j over
 * the set bits j of r), column j in bits 4j..4j+3 and c in bits 16..19.
 *
 * A matrix-core block (kind 3) applies a dense 2^k x 2^k complex matrix M
 * with FP64 MFMA.  Its coefficients are the real form
 * R = [[Re M, -Im M], [Im M, Re M]] (2^(k+1) square) packed in operand order:
 * double P[(rb * NK + ks) * 64 + l] = R[16 rb + l % 16][4 ks + l / 16] for
 * row block rb < 2^(k+1) / 16, k-step ks < NK = 2^(k+1) / 4 and lane l < 64,
 * i.e. 4^(k+1) doubles = 2 * 4^k complex128 slots.  It needs
 * tile_bits >= k + 4. */
#define QFX_TILE_WIDE 1   /* flags: the pass holds 4-5 qubit dense/sparse ops */
#define QFX_TILE_PERSISTENT 2  /* flags: persistent workgroups that prefetch
                                  the next tile during the current one (used
                                  when 2^tile_bits / block <= 16) */
#define QFX_TILE_MFMA 4   /* flags: the pass holds matrix-core blocks */
#define QFX_TILE_REG_IO 8 /* flags: persistent pass whose last op is a register
                             phase storing straight to HBM; needs
                             2^tile_bits == 16 * block, block <= 512 */
#define QFX_TILE_REG_IN 16 /* flags: persistent pass whose first op is a register
                              phase loading straight from HBM (no LDS round
                              trip before it); needs 2^tile_bits == 16 * block,
                              block <= 512; combines with QFX_TILE_REG_IO */

/* Fills the derived `lds` words of every register phase in ops[0..nops-1]
 * (host array) from its tile bits and maps; zeroes them for other kinds.
 * Call after the last change to an op and before qfx_tile_check_ops. */
int qfx_tile_prepare_ops(qfx_tile_op* ops /*host*/, int nops);

/* Host-side validation of an op list against tile width T, the block size
 * and flags it will run with (ops on 4-5 qubits need QFX_TILE_WIDE and
 * block <= 256; matrix-core blocks need QFX_TILE_MFMA; other passes at most
 * 768 coefficients), ncoef coefficients and nsrcs source rows (host arrays);
 * register phases must carry the `lds` words qfx_tile_prepare_ops gives
 * them.  The device entry point trusts its op list; the Python planner
 * checks every plan with this before upload. */
int qfx_tile_check_ops(int T, int block, int flags, const qfx_tile_op* ops /*host*/,
                       int nops, int ncoef, const int* srcs /*host*/,
                       int nsrcs);

/* One pass: 6 <= tile_bits <= 13, tile_bits <= n, block in {64, 256, 512,
 * 1024} with 2^tile_bits >= block; flags as accepted by qfx_tile_check_ops
 * for this op list.  ops/coef/srcs in device memory; coef holds the op
 * list's ncoef coefficients (indices in the ops are relative to coef).  A
 * pass without QFX_TILE_WIDE or QFX_TILE_MFMA takes at most 768
 * coefficients: it stages them in LDS.  Tile bit j is loaded from index bit
 * tile_pos[j].  The pass runs in place and stores source index bit b to
 * index bit dest[b] (NULL: the identity): dest may permute the tile
 * positions among themselves (a relabelling of the pass's qubits, the SWAP
 * gates it absorbed) and must keep every other index bit. */
int qfx_tile_pass(void* stream, void* state, int n, int tile_bits,
                  const int* tile_pos /*host*/, const int* dest /*host or NULL*/,
                  const void* ops, int nops, const void* coef, int ncoef,
                  const void* srcs, int block, int flags);

/* ---- QAOA MaxCut (quantumflow/qaoa.py) --------------------------------- */

/* Cut table C(z) = sum_e w_e [z_u != z_v] over all 2^n basis states, edges
 * summed in the given order (bitwise equal to graph_cuts,
 * quantumflow/qaoa.py:68-81).  edges: device int32 [m][2] with node labels
 * 0..n-1 (node u <-> axis u); weights: device float64 [m] or NULL for 1.0;
 * m <= 4096. */
int qfx_cut_table(void* stream, void* out, int n, const void* edges,
                  const void* weights, int m);

/* Unit-weight cut counts as uint16 (the same table as integers). */
int qfx_cut_counts(void* stream, void* out, int n, const void* edges, int m);

/* ---- batched QAOA engine (quantumflow_amd/qaoa_engine.py) -------------
 * Expectation <C> and its adjoint gradient for many MaxCut instances in one
 * persistent launch.  Replaces per instance the chain qubo_circuit ->
 * Circuit.run -> State.expectation(graph_cuts) (quantumflow/qaoa.py:22-81,
 * quantumflow/states.py:131-147) and the autograd gradient of
 * examples/qaoa_maxcut.py:53-64.  The host builds, per instance, a program
 * of tiled passes (tile bit set + stages); global pass k of all instances
 * runs between grid barriers.  Every array below is device memory. */

#define QFX_QAOA_MAX_LAYERS 16
#define QFX_QAOA_MAX_STAGES 40
#define QFX_QAOA_FORWARD 0   /* width 1: <C> */
#define QFX_QAOA_BACKWARD 1  /* width 1 + 2p: <C>, d/dbeta_l, d/dgamma_l */

typedef struct {
  int32_t n;          /* qubits (graph nodes 0..n-1, node u <-> bit n-1-u) */
  int32_t p;          /* layers */
  int32_t weighted;   /* 0: uint16 cut counts; 1: float64 cut values */
  int32_t nedges;     /* edges (phase-table size for unit weights) */
  double wsum;        /* sum of edge weights W */
  int64_t cut_off;    /* first cut-table entry in cnt_pool / val_pool */
  int64_t psi_off;    /* first amplitude of the state in psi_pool */
  int64_t lam_off;    /* first amplitude of lambda in lam_pool (backward) */
  int32_t angle_off;  /* beta[0..p-1] then gamma[0..p-1] in angles */
  int32_t slot0;      /* first partial slot of this instance */
  int32_t nslots;     /* partial slots of this instance */
  int32_t mirror;     /* 1: the instance runs on half of its state.  MaxCut
                         states are symmetric under flipping every bit, so an
                         amplitude whose top index bit (n-1) is set equals that
                         of the complementary index: psi and lambda hold the
                         2^(n-1) amplitudes with bit n-1 clear, every pass has
                         bit n-1 as its highest tile bit, its items cover only
                         the tiles whose highest non-tile bit is clear, and an
                         element with bit n-1 set is loaded from and stored to
                         the complementary index.  Values and gradients count
                         each item twice.  0: the whole state. */
} qfx_qaoa_inst;

typedef struct {
  int32_t T;          /* tile bits, <= the program's tile_bits */
  int32_t pos[13];    /* tile bit positions, ascending */
  int32_t nstage;
  int32_t flags;      /* 1 load psi, 2 load lambda, 4 store psi, 8 store lambda */
  /* {type, layer, tile-bit mask}; types: 1 init |+>, 2 cost layer, 3 RX on
   * the mask, 4 <C>, 5 lambda = C psi, 6 adjoint RX, 7 adjoint cost layer,
   * 8 cost layer then RX of the same layer (forward programs only; the
   * phases ride in the first RX chunk; mask must be non-empty), 10 / 11 RX
   * applied while the tile is loaded / stored (forward programs only; the
   * pass's first stage with flag 1 / last stage with flag 4; the bits a
   * thread holds in registers then, the 4 tile bits above log2(block), cost
   * no LDS sweep, any others run as an LDS chunk) */
  int32_t stage[QFX_QAOA_MAX_STAGES][3];
} qfx_qaoa_pass;

typedef struct {
  int32_t inst;       /* instance */
  int32_t pass;       /* pass descriptor */
  int32_t tile;       /* tile index within the pass */
  int32_t partial;    /* partial slot written by this item */
} qfx_qaoa_item;

typedef struct {
  const void* insts;      /* qfx_qaoa_inst[ninst] */
  const void* passes;     /* qfx_qaoa_pass[] */
  const void* items;      /* qfx_qaoa_item[], grouped by global pass */
  const void* kstart;     /* int32[nk + 1]: items of global pass k */
  const void* angles;     /* float64 */
  void* psi_pool;         /* complex128 */
  void* lam_pool;         /* complex128 (backward) */
  const void* cnt_pool;   /* uint16 cut counts */
  const void* val_pool;   /* float64 cut values */
  void* partials;         /* float64 [nslots][width] */
  void* out;              /* float64 [ninst][width] */
  void* sync;             /* >= 12 bytes: int32 barrier counter, this launch's
                             error flag, sticky error flag.  The owner zeroes
                             all three before the first launch; a launch
                             resets the first two and leaves the third, which
                             any grid-barrier timeout sets. */
  int32_t ninst, nk, width, mode, block, max_items_per_k;
  int32_t tile_bits;      /* largest pass tile (LDS per workgroup: 2^tile_bits
                             amplitudes, twice that backward); 0: 13 forward,
                             12 backward.  Must be what qfx_qaoa_check
                             accepted the program with. */
  int32_t pad;
  void* timing;           /* NULL, or uint64 [timing_len]: per global pass k
                             and workgroup b, entries 2 (k grid + b) and +1
                             get the device wall clock when b starts pass k
                             and when it reaches the barrier after it (entries
                             past timing_len are not written) */
  int64_t timing_len;
} qfx_qaoa_args;

/* Host-side validation of a program (host copies of the device arrays);
 * tile_bits as in qfx_qaoa_args (0: the mode's default). */
int qfx_qaoa_check(const qfx_qaoa_inst* insts, int ninst,
                   const qfx_qaoa_pass* passes, int npasses,
                   const qfx_qaoa_item* items, const int* kstart, int nk,
                   int width, int64_t psi_pool_len, int64_t lam_pool_len,
                   int64_t cnt_pool_len, int64_t val_pool_len,
                   int64_t npartial, int mode, int tile_bits);

/* Enqueue the program: a barrier reset, the persistent kernel (grid = the
 * workgroups that fit the device at once) and the per-instance reduction
 * into out.  Capturable into a HIP graph. */
int qfx_qaoa_run(void* stream, const qfx_qaoa_args* args);

/* After the stream completes: *failed = 1 if a grid barrier timed out in
 * the last launch or in any earlier launch since the owner zeroed the
 * sticky flag.  A barrier times out after 10 s of device wall-clock time. */
int qfx_qaoa_status(const void* sync, int* failed);

/* ---- MLP angle proposer kernels (fp32) ----------------------------------
 * No counterpart in the reference (which has no learned component): the
 * proposer warm-starts the optimizer of examples/qaoa_maxcut.py:37-87. */

/* C[m][n] = sum_k A[m*sam + k*sak] * B[k*sbk + n*sbn] (+ bias[n]) (relu);
 * C is row-major [M][N].  Fixed k order per element: reproducible. */
int qfx_mlp_gemm(void* stream, void* C, const void* A, const void* B, int M,
                 int N, int K, int64_t sam, int64_t sak, int64_t sbk,
                 int64_t sbn, const void* bias /*[N] or NULL*/, int relu);
/* dz = dy * (y > 0) */
int qfx_mlp_relu_bwd(void* stream, void* dz, const void* dy, const void* y,
                     int64_t len);
/* out[n] = sum_b x[b][n]  (bias gradient) */
int qfx_mlp_colsum(void* stream, void* out, const void* x, int B, int N);
/* Bytes of workspace qfx_mlp_mse needs. */
int64_t qfx_mlp_mse_workspace(int64_t len);
/* loss[0] = mean (y - t)^2 ; dy = 2 (y - t) / len */
int qfx_mlp_mse(void* stream, void* dy, void* loss, const void* y,
                const void* t, int64_t len, void* workspace);
/* Adam (Kingma & Ba 2015) with bias correction for step >= 1 and optional
 * L2 weight decay added to the gradient. */
int qfx_mlp_adam(void* stream, void* params, const void* grads, void* m,
                 void* v, int64_t len, float lr, float beta1, float beta2,
                 float eps, int step, float weight_decay);

#ifdef __cplusplus
}
#endif

#endif /* QFX_H */

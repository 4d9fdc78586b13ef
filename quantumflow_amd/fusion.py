"""Gate fusion and LDS-tiled execution plans.

The plain path of Circuit.run (quantumflow/circuits.py:87-99) streams the
whole state through HBM once per gate.  A TilePlan instead groups the gate
list into *passes*: each pass picks a set S of T index bits (the lowest L
bits always, for coalescing, plus the qubits of the gates it absorbs) and
the tile kernel (csrc/tile.hip) applies all of the pass's gates to every
2^T-amplitude tile in LDS, so a pass costs one HBM read + write of the
state however many gates it holds.

Planning (host, once per circuit):
  1. ``plan_passes``: greedy in circuit order.  A gate joins the current
     pass if its bits fit in S (growing S up to T bits) and it does not
     touch a qubit of a gate already deferred to a later pass (so it
     commutes with every gate it overtakes); otherwise it is deferred and
     its qubits are blocked for the rest of the pass.
  2. Within a pass, by ``mode``:
     'reg'   ``plan_phases`` groups gates into register phases of 4 tile
             bits (one LDS sweep each, gates applied in registers);
     'mfma'  ``plan_blocks`` merges all gates into dense blocks of 3 to
             ``fuse_width`` (<= 5) qubits, applied by FP64 MFMA;
     'lds'   ``fuse_ops`` merges runs of consecutive gates into dense blocks
             of <= ``fuse_width`` qubits when a cost model (LDS sweeps + FP64
             FMA work) prefers it; one LDS op each.
     A block's matrix is the product of its gates' matrices embedded in the
     block (Circuit.asgate semantics, quantumflow/circuits.py:111-118).
  3. The op tables of all passes go to the device once; running the plan is
     one kernel launch per pass.
"""

import threading
from collections import OrderedDict, defaultdict

import numpy as np
import torch

from . import _native
from . import backend as bk
from .states import State

OP_DENSE, OP_SPARSE, OP_RPHASE, OP_MFMA = 0, 1, 2, 3
SUB_D1, SUB_DIAG, SUB_CX, SUB_D1R, SUB_D1X, SUB_DIAGV = 1, 4, 5, 7, 8, 9
# record flags of sub-ops that read a bit a lane flip may have hit (2x2
# matrices: 8 coefficients, m then X m X; CX: controls through the flips)
D1_FLIP, CX_FLIP = 4, 1 << 10
# a record's type sits in the low bits of its word, a lane flip riding on
# the record in the bits above
REC_TYPE = 255
MAX_FLIP_SEL = 2            # selectors of one lane flip
MAX_SEL_BITS = 3          # index bits outside the tile a register diagonal may read
REG_BITS = 4
MAX_TILE_BITS = 13
MIN_TILE_BITS = 6
TILE_WIDE = 1
TILE_PERSISTENT = 2
TILE_MFMA = 4
TILE_REG_IO = 8
TILE_REG_IN = 16
MFMA_MIN_BITS = 3


class KernelTimer:
    """HIP-event timing of individual launches on torch's current stream
    (the stream every qfx call is enqueued on)."""

    def __init__(self):
        self.events = defaultdict(list)

    def start(self, name):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return name, e

    def stop(self, tok):
        name, e0 = tok
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.events[name].append((e0, e1))

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for name, pairs in self.events.items():
            ts = [a.elapsed_time(b) for a, b in pairs]
            out[name] = (sum(ts), len(ts))
        return out

    def dominant(self):
        """(kernel name, average ms per launch, launches) of the kernel class
        with the largest total time."""
        s = self.summary()
        if not s:
            return None, float('nan'), 0
        name = max(s, key=lambda k: s[k][0])
        tot, cnt = s[name]
        return name, tot / cnt, cnt


def kernel_class(kind):
    return '%s_k%d' % (kind.kind, kind.k)


def apply_gates_inplace(gates, qubits, work, prof=None):
    """Unfused path on a flat device state ``work``: one in-place kernel per
    gate (diagonal/monomial gates touch only the rows they change)."""
    qubits = tuple(qubits)
    n = len(qubits)
    for g in gates:
        bits = [n - 1 - qubits.index(q) for q in g.qubits]
        kind = g.kind
        if prof is not None:
            tok = prof.start(kernel_class(kind))
        bk.apply_gate(work, work, n, bits, kind, inplace_sparse=True)
        if prof is not None:
            prof.stop(tok)
    return work


# ---------------------------------------------------------------------------
# planning
# ---------------------------------------------------------------------------

def gate_types(kind, bits):
    """How a gate acts on each of its bits, for commutation: 'Z' where it
    commutes with Z on that bit (diagonal gates; controls of controlled-X
    gates), 'X' where it acts as a function of X alone (X, RX; targets of
    controlled-X gates), 'G' otherwise.  Two gates whose shared bits all
    carry the same letter, Z or X, commute."""
    if kind.kind == bk.GateKind.DIAGONAL:
        return {b: 'Z' for b in bits}
    pat = exchange_pattern(kind) if kind.kind == bk.GateKind.MONOMIAL else None
    if pat is not None and pat[0] == 'cx' and kind.phase is None:
        t = {b: 'Z' for b in bits}
        t[bits[pat[1]]] = 'X'
        return t
    if kind.k == 1:
        m = kind.matrix
        if m[0, 0] == m[1, 1] and m[0, 1] == m[1, 0]:
            return {bits[0]: 'X'}
    return {b: 'G' for b in bits}


def _absorb(bits_list, pending, S, ok=None, free=None, types=None, need=None):
    """Items (indices into bits_list, in order) a set of bits S takes: an item
    joins if it is allowed, the bits it needs lie in S (all of its bits, or
    ``need[i]``: a controlled X needs only its target; a ``free`` diagonal
    gate needs none) and it commutes with every item left behind before it:
    it shares no bit with them or, with ``types`` (gate_types per item),
    every shared bit carries the same 'Z' or 'X' in both."""
    blocked = {}
    taken = []
    for i in pending:
        b = bits_list[i]
        nb = b if need is None else need[i]
        if types is None:
            clear = not any(x in blocked for x in b)
        else:
            ti = types[i]
            clear = all(x not in blocked or (ti[x] != 'G' and blocked[x] <= {ti[x]}) for x in b)
        if (clear and (ok is None or ok[i])
                and (S.issuperset(nb) or (free is not None and free[i]))):
            taken.append(i)
            continue
        for x in b:
            blocked.setdefault(x, set()).add('G' if types is None else types[i][x])
    return taken


def _grow(bits_list, pending, S, width, ok=None, cap=40, free=None, types=None, need=None):
    """Grow the bit set S to at most ``width`` bits, one candidate at a time:
    the candidates are the needed bit sets of the first ``cap`` distinct
    items not yet taken (in circuit order), and the one that takes the most
    further items per added bit wins.  Returns (S, items taken)."""
    cur = _absorb(bits_list, pending, S, ok, free, types, need)
    while len(S) < width:
        taken = set(cur)
        best = None
        seen = set()
        for i in pending:
            if i in taken or (ok is not None and not ok[i]) or (free is not None and free[i]):
                continue
            new = frozenset(bits_list[i] if need is None else need[i]) - S
            if not new or new in seen or len(S) + len(new) > width:
                continue
            seen.add(new)
            got = _absorb(bits_list, pending, S | new, ok, free, types, need)
            gain = (len(got) - len(cur)) / len(new)
            if best is None or gain > best[0]:
                best = (gain, new, got)
            if len(seen) >= cap:
                break
        if best is None or best[0] <= 0:
            break
        S = S | best[1]
        cur = best[2]
    return S, cur


def plan_passes(bits_list, n, T, L, tileable=None, strategy='gain', free=None, types=None,
                need=None):
    """Group gates (given by their index-bit tuples, in circuit order) into
    passes.  Returns a list of (sorted tile bits, [gate indices]); a gate
    that is not tileable forms a pass of its own with tile bits None.
    ``free[i]`` marks diagonal gates that a pass can apply whatever its tile
    bits (register phases select their phases by the tile's other index
    bits): they join the first pass that reaches them and add no tile bits.
    ``types`` (gate_types per gate) lets a gate overtake the commuting gates
    a pass leaves behind even where they share bits.  ``need[i]`` (default:
    all of gate i's bits) are the bits gate i needs in the tile: a
    controlled X whose controls may lie outside it needs only its target.

    strategy 'gain' grows each pass's tile set by ``_grow`` (the bits that
    take the most gates per added bit); 'greedy' takes gates first come,
    first served (28-qubit depth-40 benchmark circuit at T=12: 40 passes
    against 51)."""
    pending = list(range(len(bits_list)))
    passes = []
    low = set(range(min(L, n)))
    while pending:
        first = pending[0]
        if tileable is not None and not tileable[first]:
            passes.append((None, [first]))
            pending = pending[1:]
            continue
        if strategy == 'gain':
            S, _ = _grow(bits_list, pending, set(low), T, tileable, free=free, types=types,
                         need=need)
        else:
            S = set(low)
            blocked = set()
            for gi in pending:
                bs = set(bits_list[gi])
                if (bs & blocked) or (tileable is not None and not tileable[gi]):
                    blocked |= bs
                    continue
                if len(S | bs) <= T:
                    S |= bs
                else:
                    blocked |= bs
        # fill the tile with the lowest unused bits: longer contiguous runs
        b = 0
        while len(S) < T:
            if b not in S:
                S.add(b)
            b += 1
        taken = _absorb(bits_list, pending, S, tileable, free, types, need)
        passes.append((sorted(S), taken))
        done = set(taken)
        pending = [g for g in pending if g not in done]
    return passes


def embed(matrix, pos, nq):
    """2^nq x 2^nq operator of a k-qubit gate whose qubit j sits at block
    position pos[j] (block qubit 0 = most significant)."""
    k = len(pos)
    g = np.asarray(matrix, dtype=np.complex128).reshape([2] * (2 * k))
    t = np.eye(2 ** nq, dtype=np.complex128).reshape([2] * nq + [2 ** nq])
    out = np.tensordot(g, t, axes=(list(range(k, 2 * k)), list(pos)))
    out = np.moveaxis(out, list(range(k)), list(pos))
    return out.reshape(2 ** nq, 2 ** nq)


class SelDiagonal:
    """A diagonal gate of a pass, part of whose qubits lie outside the tile:
    a diagonal on its in-tile bits whose phases depend on the tile's other
    index bits ``sel`` (global index bits).  ``tables[v]`` is the diagonal on
    the in-tile bits (gate order) when bit j of v is the value of index bit
    sel[j].  Duck-types a diagonal GateKind."""
    kind = 'diagonal'
    src = phase = matrix = None

    def __init__(self, kind, bits, tile_bits):
        ins = [j for j, b in enumerate(bits) if b in tile_bits]
        outs = [j for j, b in enumerate(bits) if b not in tile_bits]
        self.in_bits = [bits[j] for j in ins]
        self.sel = [bits[j] for j in outs]
        self.k = len(ins)
        k = len(bits)
        tables = np.empty((2 ** len(outs), 2 ** len(ins)), dtype=np.complex128)
        for v in range(2 ** len(outs)):
            for x in range(2 ** len(ins)):
                r = 0
                for a, j in enumerate(ins):
                    r |= ((x >> (len(ins) - 1 - a)) & 1) << (k - 1 - j)
                for a, j in enumerate(outs):
                    r |= ((v >> a) & 1) << (k - 1 - j)
                tables[v, x] = kind.diag[r]
        self.tables = tables
        self.diag = None


class SelCX:
    """A controlled X of a pass some of whose controls lie outside the tile:
    X on the target, controlled by its in-tile controls (a monomial on the
    in-tile bits, target first) and by the index bits ``sel`` outside the
    tile, which fix per tile whether it acts.  All controls are on |1>.
    Duck-types a monomial GateKind."""
    kind = 'monomial'
    phase = matrix = diag = None

    def __init__(self, target, in_ctrls, sel):
        self.in_bits = [target] + list(in_ctrls)
        self.sel = list(sel)
        k = self.k = len(self.in_bits)
        # output row r takes input row src[r]: flip the target (gate qubit 0,
        # the most significant) where every in-tile control is 1
        ctl = (1 << (k - 1)) - 1
        self.src = np.array([r ^ (1 << (k - 1)) if (r & ctl) == ctl else r
                             for r in range(2 ** k)])


def cx_parts(op_or_kind, bits):
    """(target bit, control bits) of a controlled X without phases acting on
    ``bits`` (gate-qubit order), else None."""
    kd = op_or_kind
    if kd.kind != bk.GateKind.MONOMIAL or kd.phase is not None:
        return None
    pat = exchange_pattern(kd)
    if pat is None or pat[0] != 'cx':
        return None
    return bits[pat[1]], [bits[c] for c in pat[2]]


class Op:
    """A gate inside a pass: bits in gate-qubit order and its GateKind (or
    a dense merged block)."""
    __slots__ = ('bits', 'kind', 'matrix', 'ngates')

    def __init__(self, bits, kind=None, matrix=None, ngates=1):
        self.bits = list(bits)
        self.kind = kind
        self.matrix = matrix
        self.ngates = ngates

    def dense_matrix(self):
        if self.matrix is not None:
            return self.matrix
        return self.kind.matrix

    def is_dense(self):
        return self.matrix is not None or self.kind.kind == bk.GateKind.DENSE

    def active_rows(self):
        kd = self.kind
        d = 2 ** kd.k
        if isinstance(kd, SelDiagonal):
            return d
        if kd.kind == bk.GateKind.DIAGONAL:
            return int(np.sum(kd.diag != 1.0))
        src = kd.src
        ph = kd.phase if kd.phase is not None else np.ones(d)
        return int(np.sum((src != np.arange(d)) | (ph != 1.0)))


# cost of one op per tile, in units of one full LDS read+write sweep
FLOP_WEIGHT = 0.15


def op_cost(op):
    k = len(op.bits)
    if op.is_dense():
        return 1.0 + FLOP_WEIGHT * (2 ** k)
    return op.active_rows() / float(2 ** k)


def fuse_ops(ops, width):
    """Merge runs of consecutive ops into dense blocks of <= width qubits
    when the cost model prefers the block."""
    if width < 2:
        return ops
    out = []
    cur = None
    for op in ops:
        if cur is None:
            cur = op
            continue
        q = list(cur.bits) + [b for b in op.bits if b not in cur.bits]
        if len(q) <= width:
            merged_cost = 1.0 + FLOP_WEIGHT * (2 ** len(q))
            if merged_cost <= op_cost(cur) + op_cost(op):
                m1 = embed(cur.dense_matrix(), [q.index(b) for b in cur.bits], len(q))
                m2 = embed(op.dense_matrix(), [q.index(b) for b in op.bits], len(q))
                cur = Op(q, None, m2 @ m1, cur.ngates + op.ngates)
                continue
        out.append(cur)
        cur = op
    if cur is not None:
        out.append(cur)
    return out


def _encode_op(op, pos_index, coefs, srcs):
    """Append op's coefficients/sources and return its qfx_tile_op."""
    o = _native.QfxTileOp()
    k = len(op.bits)
    t = [pos_index[b] for b in op.bits]
    o.k = k
    for j in range(5):
        o.t[j] = t[j] if j < k else 0
    st = sorted(t)
    for j in range(5):
        o.sorted[j] = st[j] if j < k else 0
    o.coef = len(coefs)
    if op.is_dense():
        o.kind = OP_DENSE
        coefs.extend(np.asarray(op.dense_matrix(), dtype=np.complex128).reshape(-1))
        o.src = 0
        o.active = o.phase = 0
    else:
        kd = op.kind
        d = 2 ** k
        o.kind = OP_SPARSE
        if kd.kind == bk.GateKind.DIAGONAL:
            src = np.arange(d)
            ph = kd.diag
        else:
            src = kd.src
            ph = kd.phase if kd.phase is not None else np.ones(d, dtype=np.complex128)
        coefs.extend(np.asarray(ph, dtype=np.complex128))
        o.src = len(srcs)
        srcs.extend(int(x) for x in src)
        active = 0
        phase = 0
        for r in range(d):
            if ph[r] != 1.0:
                phase |= 1 << r
            if ph[r] != 1.0 or src[r] != r:
                active |= 1 << r
        o.active, o.phase = active, phase
    return o



def exchange_pattern(kind):
    """Describe a monomial gate's row permutation as a controlled X or a
    controlled SWAP of its qubits (controls on |1>): ('cx', t, controls) or
    ('swap', a, b, controls) in gate-qubit numbers, or None."""
    if kind.kind != bk.GateKind.MONOMIAL:
        return None
    k = kind.k
    src = [int(x) for x in kind.src]
    q = list(range(k))

    def bit(r, j):                      # value of gate qubit j in row r
        return (r >> (k - 1 - j)) & 1

    for t in q:
        others = [j for j in q if j != t]
        for m in range(1 << len(others)):
            ctl = [others[i] for i in range(len(others)) if (m >> i) & 1]
            want = [r ^ (1 << (k - 1 - t)) if all(bit(r, c) for c in ctl) else r
                    for r in range(2 ** k)]
            if want == src:
                return ('cx', t, ctl)
    for a in q:
        for b in q:
            if b <= a:
                continue
            others = [j for j in q if j not in (a, b)]
            for m in range(1 << len(others)):
                ctl = [others[i] for i in range(len(others)) if (m >> i) & 1]
                want = []
                for r in range(2 ** k):
                    if all(bit(r, c) for c in ctl) and bit(r, a) != bit(r, b):
                        r = r ^ (1 << (k - 1 - a)) ^ (1 << (k - 1 - b))
                    want.append(r)
                if want == src:
                    return ('swap', a, b, ctl)
    return None


def phaseable(op):
    """Can the op run inside a register phase?"""
    if op.matrix is not None:
        return len(op.bits) == 1
    kd = op.kind
    if kd.kind == bk.GateKind.DIAGONAL:
        return kd.k <= REG_BITS
    if kd.kind == bk.GateKind.MONOMIAL and exchange_pattern(kd) is not None:
        return kd.k <= REG_BITS
    # dense 2-qubit gates run as LDS ops (no 4x4 matrices in registers)
    return kd.k == 1


def is_diagonal_op(op):
    return op.matrix is None and op.kind.kind == bk.GateKind.DIAGONAL


def plan_phases(ops, width=REG_BITS, free_diagonals=False):
    """Group a pass's ops (in order) into register phases of <= width bits
    and standalone LDS ops, with the commuting-overtake rule of
    plan_passes and its bit-set growth.  With ``free_diagonals`` a diagonal
    op joins the first phase that reaches it whatever that phase's register
    bits (its other bits select its phases per lane group), and a
    controlled X needs only its target in registers (its other controls
    become selectors of the same kind).  Returns [('phase', bits, ops) |
    ('lds', op)]."""
    bits_list = [tuple(op.bits) for op in ops]
    ok = [phaseable(op) for op in ops]
    free = need = None
    if free_diagonals:
        free = [is_diagonal_op(op)
                and len(op.bits) + len(getattr(op.kind, 'sel', ())) <= MAX_SEL_BITS
                for op in ops]
        need = []
        for op, b in zip(ops, bits_list):
            cx = cx_parts(op.kind, b) if op.matrix is None else None
            nsel = len(getattr(op.kind, 'sel', ())) + (len(cx[1]) if cx else 0)
            need.append((cx[0],) if cx and nsel <= MAX_SEL_BITS else b)
    pending = list(range(len(ops)))
    items = []
    while pending:
        first = pending[0]
        if not ok[first]:
            items.append(('lds', ops[first]))
            pending = pending[1:]
            continue
        R, taken = _grow(bits_list, pending, set(), width, ok, free=free, need=need)
        items.append(('phase', R, [ops[i] for i in taken]))
        done = set(taken)
        pending = [i for i in pending if i not in done]
    return items


def _diag_tables(op, reg_of, pos_index):
    """Selector codes and [2^m, 16] logical-slot phase tables of a diagonal
    op in a register phase: gate qubits on the phase's register bits index
    the 16 slots; the others are selectors, code 64 + t for tile bit t (the
    lane group's bit) or the state-index bit b itself for bits outside the
    tile (SelDiagonal.sel)."""
    kd = op.kind
    if isinstance(kd, SelDiagonal):
        k_in = len(op.bits)
        gsel = list(kd.sel)

        def value(x_in, v_glob):
            return kd.tables[v_glob][x_in]
    else:
        k_in = len(op.bits)
        gsel = []

        def value(x_in, v_glob):
            return kd.diag[x_in]
    regs, lane = [], []
    for j, b in enumerate(op.bits):
        t = pos_index[b]
        if t in reg_of:
            regs.append((j, reg_of[t]))
        else:
            lane.append((j, t))
    sel = [64 + t for _, t in lane] + gsel
    m = len(sel)
    tab = np.empty((2 ** m, 2 ** REG_BITS), dtype=np.complex128)
    for v in range(2 ** m):
        for r in range(2 ** REG_BITS):
            x_in = 0
            for j, a in regs:
                x_in |= ((r >> a) & 1) << (k_in - 1 - j)
            for i, (j, _) in enumerate(lane):
                x_in |= ((v >> i) & 1) << (k_in - 1 - j)
            tab[v, r] = value(x_in, v >> len(lane))
    return sel, tab


def _slot_phases(diag, p):
    """16 register-slot phases of a diagonal on register bits p (gate MSB
    first)."""
    k = len(p)
    ph = np.empty(2 ** REG_BITS, dtype=np.complex128)
    for r in range(2 ** REG_BITS):
        idx = 0
        for j in range(k):
            idx |= ((r >> p[j]) & 1) << (k - 1 - j)
        ph[r] = diag[idx]
    return ph


def d1_variant(m):
    """Cheapest 1-qubit sub-op that applies 2x2 matrix m exactly: real (RY,
    H: half the FMAs), real diagonal with imaginary off-diagonal (RX: half
    the FMAs) or general."""
    m = np.asarray(m, dtype=np.complex128)
    if not np.any(m.imag):
        return SUB_D1R
    if not np.any(m.real[[0, 1], [1, 0]]) and not np.any(m.imag[[0, 1], [0, 1]]):
        return SUB_D1X
    return SUB_D1


def _encode_phase(R, members, pos_index, T, coefs, subs, flips=True, store_cx=True,
                  merge_records=True):
    """qfx_tile_op of a register phase over tile bits R plus its sub-ops.

    Logical bit A of the phase is tile bit regs[A]; register i starts out
    holding logical index i.  Uncontrolled exchanges (SWAP, X and the
    permutation part of ISWAP, Y) are not executed: they relabel which
    logical index each register holds, phi(i) = P(i) xor c with P a bit
    permutation (q[A] = register bit of logical bit A), and later sub-ops
    are translated through phi.  The op's store map writes register i to
    the slot of logical phi(i), and a leading run of SWAP/X/CNOT gates and
    diagonals is folded into the load map.  Both maps are affine on the
    4-bit logical index and are stored as matrix | offset << 16 (column j
    = image of bit j): ``active`` the store map, ``phase`` the load map.

    With ``flips``, a controlled X whose controls are all selectors (no
    register bit) becomes a lane flip riding on a later record: it moves no
    amplitudes.  It may ride past records that do not read its bit (it
    commutes with them), one flip per record.  Later diagonals name the
    flip-mask bits they read; later 2x2 matrices on its bit and CX gates
    controlled by it carry a flag."""
    tl = sorted(pos_index[b] for b in R)
    # diagonals whose bits all lie in R may fold into the load map; the others
    # (bits outside R: lane-group or outside-tile selectors) may not
    def in_regs(op):
        return all(b in R for b in op.bits)
    for j in range(T - 1, -1, -1):          # pad to 4 register bits with high tile
        if len(tl) >= REG_BITS:             # bits: lanes keep the low ones, whose
            break                           # amplitudes are contiguous in HBM
        if j not in tl:
            tl.append(j)
    regs = tl[:REG_BITS]                    # logical bit A <-> tile bit regs[A]
    reg_of = {t: j for j, t in enumerate(regs)}
    o = _native.QfxTileOp()
    o.kind = OP_RPHASE
    o.k = REG_BITS
    for j in range(5):
        o.t[j] = regs[j] if j < REG_BITS else 0
    st = sorted(regs)
    for j in range(5):
        o.sorted[j] = st[j] if j < REG_BITS else 0
    while len(subs) % 4:                    # records are read as aligned int4
        subs.append(0)
    o.src = len(subs)
    out = []                                # pending sub-ops: [type, code, matrix/phases]
    q = list(range(REG_BITS))               # register bit holding logical bit A
    flip = [0]                              # logical bits flipped (c)

    def phi(i):
        x = 0
        for a in range(REG_BITS):
            x |= ((i >> q[a]) & 1) << a
        return x ^ flip[0]

    # Diagonal factors wait in `pending` (register space) and are emitted as
    # one SUB_DIAG only before a 1-qubit sub-op on a bit they depend on, or
    # at the end: an exchange E moves past a pending diagonal D as
    # E D = D' E with D'[i] = D[E(i)].  A pending diagonal is a table
    # [2^m, 16]: row v holds the phases when bit j of v is the value of the
    # index bit sel[j] outside the tile (m <= MAX_SEL_BITS; SUB_DIAGV).
    pending = [None]
    regs16 = np.arange(2 ** REG_BITS)

    def expand(tab, sel, union):
        rows = [sum(((v >> union.index(b)) & 1) << j for j, b in enumerate(sel))
                for v in range(2 ** len(union))]
        return tab[rows]

    def push_diag(ph_log, sel=(), tab_log=None):
        tab = np.asarray(ph_log)[None, :] if tab_log is None else np.asarray(tab_log)
        tab = tab[:, [phi(i) for i in range(2 ** REG_BITS)]]
        sel = list(sel)
        if pending[0] is not None:
            psel, ptab = pending[0]
            union = psel + [b for b in sel if b not in psel]
            if len(union) <= MAX_SEL_BITS:
                pending[0] = (union, expand(ptab, psel, union) * expand(tab, sel, union))
                return
            flush()
        pending[0] = (sel, tab)

    def flush(bit=None):
        d = pending[0]
        if d is None:
            return
        sel, tab = d
        if bit is not None and np.array_equal(tab, tab[:, regs16 ^ (1 << bit)]):
            return
        out.append([SUB_DIAGV if sel else SUB_DIAG, sel, tab])
        pending[0] = None

    def push_cx_sel(t, ctrls, sels):
        """X on logical bit t controlled by logical bits ctrls and by the
        selector codes sels (lane-group tile bits 64 + t, index bits outside
        the tile), all on |1>.  Whether it acts varies across lanes and
        tiles, so a pending diagonal is emitted first."""
        flush()
        cm = cv = 0
        for b in ctrls:
            cm |= 1 << q[b]
            cv |= (1 ^ ((flip[0] >> b) & 1)) << q[b]
        w = len(sels)
        for j, c in enumerate(sels):
            w |= (c << (2 + 7 * j)) | (1 << (23 + j))
        out.append([SUB_CX, q[t] | (cm << 2) | (cv << 6), w])

    def push_cx(t, ctrls):
        if not ctrls:                       # X: relabel only
            flip[0] ^= 1 << t
            return
        cm = cv = 0
        for b in ctrls:
            cm |= 1 << q[b]
            cv |= (1 ^ ((flip[0] >> b) & 1)) << q[b]
        if pending[0] is not None:
            e = np.where((regs16 & cm) == cv, regs16 ^ (1 << q[t]), regs16)
            pending[0] = (pending[0][0], pending[0][1][:, e])
        out.append([SUB_CX, q[t] | (cm << 2) | (cv << 6), None])

    # Leading exchanges with at most one control (SWAP, X, CNOT: affine maps
    # of the logical index) and diagonals are folded into the load map:
    # register i loads logical slot E^-1(i), so the registers start out
    # holding E psi.  Diagonals met on the way are carried past E.
    k0 = 0
    for op in members:
        kd = op.kind
        if isinstance(kd, (SelDiagonal, SelCX)) or not in_regs(op):
            break
        p = [reg_of[pos_index[b]] for b in op.bits]
        if op.matrix is None and kd.kind == bk.GateKind.DIAGONAL:
            k0 += 1
            continue
        pat = exchange_pattern(kd) if op.matrix is None else None
        if pat is None or len(pat[-1]) > 1 or (pat[0] == 'swap' and pat[-1]):
            break
        k0 += 1
    e_map = list(range(2 ** REG_BITS))      # E as a table, logical -> logical
    lead_diag = np.ones(2 ** REG_BITS, dtype=np.complex128)
    for op in members[:k0]:
        p = [reg_of[pos_index[b]] for b in op.bits]
        kd = op.kind
        if kd.kind == bk.GateKind.DIAGONAL:
            # D applied after E so far: multiply by D at the current index
            lead_diag = lead_diag * _slot_phases(kd.diag, p)
            continue
        g = _gate_perm(kd, p)
        e_map = [g[x] for x in e_map]       # x -> g(E(x))
        # a diagonal before g becomes D'(y) = D(g^-1 y) after it
        ginv = np.argsort(g)
        lead_diag = lead_diag[ginv]
        if kd.phase is not None:
            lead_diag = lead_diag * _slot_phases(kd.phase, p)
    load_inv = np.argsort(e_map)            # register i loads logical E^-1(i)
    if np.any(lead_diag != 1.0):
        push_diag(lead_diag)

    for op in members[k0:]:
        kd = op.kind
        if isinstance(kd, SelDiagonal) or (is_diagonal_op(op) and not in_regs(op)):
            sel, tab = _diag_tables(op, reg_of, pos_index)
            push_diag(None, sel, tab)
            continue
        cx = cx_parts(kd, op.bits) if op.matrix is None else None
        if cx is not None and (isinstance(kd, SelCX) or not in_regs(op)):
            # controls outside the registers: lane-group bits and, for a
            # SelCX, index bits outside the tile
            t, ctl = cx
            rc = [reg_of[pos_index[c]] for c in ctl if pos_index[c] in reg_of]
            sels = [64 + pos_index[c] for c in ctl if pos_index[c] not in reg_of]
            sels += list(getattr(kd, 'sel', ()))
            push_cx_sel(reg_of[pos_index[t]], rc, sels)
            continue
        p = [reg_of[pos_index[b]] for b in op.bits]        # logical bit per gate qubit
        if op.matrix is None and kd.kind == bk.GateKind.DIAGONAL:
            push_diag(_slot_phases(kd.diag, p))
            continue
        pat = exchange_pattern(kd) if op.matrix is None else None
        if pat is not None:
            if pat[0] == 'cx':
                push_cx(p[pat[1]], [p[c] for c in pat[2]])
            else:
                a, b = p[pat[1]], p[pat[2]]
                ctrls = [p[c] for c in pat[3]]
                if not ctrls:               # SWAP: relabel only
                    q[a], q[b] = q[b], q[a]
                    if ((flip[0] >> a) ^ (flip[0] >> b)) & 1:
                        flip[0] ^= (1 << a) | (1 << b)
                else:                       # controlled SWAP as three exchanges
                    push_cx(a, [b])
                    push_cx(b, ctrls + [a])
                    push_cx(a, [b])
            if kd.phase is not None:
                push_diag(_slot_phases(kd.phase, p))
            continue
        m = np.asarray(op.dense_matrix(), dtype=np.complex128)     # 1-qubit
        if (flip[0] >> p[0]) & 1:          # register pair holds (1, 0): X m X
            m = m[::-1, ::-1].copy()
        j = q[p[0]]
        flush(j)
        if out and out[-1][0] == SUB_D1 and out[-1][1] == j:
            out[-1][2] = m @ out[-1][2]                     # same bit: one 2x2
            continue
        out.append([SUB_D1, j, m])
    flush()

    # A 2x2 matrix on register bit j moves forward to the next one on j
    # when nothing between reads or writes bit j, and the two become one
    # record (their product).
    def deps(tab):                          # register bits a table depends on
        return sum(1 << j for j in range(REG_BITS)
                   if not np.array_equal(tab, tab[:, regs16 ^ (1 << j)]))

    def touches(entry, j):
        typ, code, data = entry
        if typ in (SUB_DIAG, SUB_DIAGV):
            return not np.array_equal(data, data[:, regs16 ^ (1 << j)])
        if typ == SUB_CX:
            return (code & 3) == j or (code >> (2 + j)) & 1
        return code == j

    def merge_matrices():
        changed = False
        k = 0
        while k < len(out):
            if out[k][0] == SUB_D1:
                j = out[k][1]
                for k2 in range(k + 1, len(out)):
                    if out[k2][0] == SUB_D1 and out[k2][1] == j:
                        if isinstance(out[k2][2], tuple) or isinstance(out[k][2], tuple):
                            break
                        out[k2] = [SUB_D1, j, out[k2][2] @ out[k][2]]
                        del out[k]
                        changed = True
                        k -= 1
                        break
                    if touches(out[k2], j):
                        break
            k += 1
        return changed

    # A diagonal moves forward to the next diagonal past 2x2 matrices on
    # bits it does not read and exchanges (an exchange E without selector
    # controls re-indexes it, D'(r) = D(E(r)); one with selector controls
    # must leave its bits alone), and the two become one record when their
    # selectors fit one table.
    def merge_diagonals():
        changed = False
        k = 0
        while k < len(out):
            if out[k][0] in (SUB_DIAG, SUB_DIAGV):
                sel, tab = out[k][1], out[k][2]
                for k2 in range(k + 1, len(out)):
                    t2, c2, d2 = out[k2]
                    if t2 in (SUB_DIAG, SUB_DIAGV):
                        union = list(sel) + [b for b in c2 if b not in sel]
                        if len(union) <= MAX_SEL_BITS:
                            prod = expand(tab, list(sel), union) * expand(d2, list(c2), union)
                            out[k2] = [SUB_DIAGV if union else SUB_DIAG, union, prod]
                            del out[k]
                            changed = True
                            k -= 1
                        break
                    if t2 == SUB_CX:
                        if d2:
                            if touches([SUB_DIAG, sel, tab], c2 & 3):
                                break
                            continue
                        cm, cv = (c2 >> 2) & 15, (c2 >> 6) & 15
                        tab = tab[:, np.where((regs16 & cm) == cv, regs16 ^ (1 << (c2 & 3)),
                                              regs16)]
                        continue
                    if touches([SUB_DIAG, sel, tab], c2):
                        break
            k += 1
        return changed

    # each merge can open the way for the other kind
    while merge_records and merge_matrices() | merge_diagonals():
        pass

    # A diagonal that reads at most register bit j folds into the 2x2 matrix
    # on j that it reaches past records leaving bit j alone: per selector
    # value v the record then holds m diag(d_v) (diagonal first) or
    # diag(d_v) m; a D1 entry with selectors carries (sel, matrices).
    def fold_diagonals():
        k = 0
        while k < len(out):
            typ, sel, tab = out[k]
            dm = deps(tab) if typ in (SUB_DIAG, SUB_DIAGV) else None
            if dm is None or dm & (dm - 1):
                k += 1
                continue
            j = dm.bit_length() - 1 if dm else None
            for step in (1, -1):
                k2 = k + step
                while 0 <= k2 < len(out):
                    t2, c2, d2 = out[k2]
                    if t2 == SUB_D1 and (j is None or c2 == j) and not isinstance(d2, tuple):
                        break
                    if t2 in (SUB_DIAG, SUB_DIAGV) or (j is not None and touches(out[k2], j)):
                        k2 = -1
                        break
                    k2 += step
                if 0 <= k2 < len(out):
                    break
            else:
                k += 1
                continue
            bit = out[k2][1]
            m = out[k2][2]
            mats = []
            for v in range(tab.shape[0]):
                d = np.diag([tab[v][0], tab[v][1 << bit]])
                mats.append(m @ d if step == 1 else d @ m)
            out[k2] = [SUB_D1, bit, (list(sel), np.asarray(mats)) if sel else mats[0]]
            del out[k]
    if merge_records:
        fold_diagonals()

    # A CNOT with its control and target in registers that commutes with
    # every sub-op after it in the phase moves to the phase's end, where
    # the affine store map absorbs it (slot r then stores where E(r) does):
    # no exchange record.  Diagonals after it take it into their tables.
    def commutes(entry, c, t):
        typ, code, data = entry
        if typ in (SUB_DIAG, SUB_DIAGV):
            return True                     # with its table re-indexed
        if typ == SUB_CX:
            return (code & 3) != c and not (code >> (2 + t)) & 1
        m = data                            # 2x2 on register bit `code`
        if isinstance(m, tuple):            # one matrix per selector value
            return code not in (c, t)
        if code == t:                       # X-symmetric: RX-type, identity
            return m[0, 0] == m[1, 1] and m[0, 1] == m[1, 0]
        if code == c:
            return m[0, 1] == 0 and m[1, 0] == 0
        return True

    deferred = {}                           # position -> register-slot permutation E
    if store_cx:
        for k, (typ, code, data) in enumerate(out):
            if typ != SUB_CX:
                continue
            cm = (code >> 2) & 15
            if data or bin(cm).count('1') != 1:
                continue
            c, t = cm.bit_length() - 1, code & 3
            # everything after it, deferred exchanges included, commutes
            if all(commutes(e, c, t) for e in out[k + 1:]):
                cv = (code >> 6) & 15
                deferred[k] = np.where((regs16 & cm) == cv, regs16 ^ (1 << t), regs16)
        # the others that commute with every sub-op before them move to the
        # phase's start, where the load map absorbs them (register s loads
        # what register E(s) would: E is an involution)
        front = {}
        for k, (typ, code, data) in enumerate(out):
            if typ != SUB_CX or k in deferred:
                continue
            cm = (code >> 2) & 15
            if data or bin(cm).count('1') != 1:
                continue
            c, t = cm.bit_length() - 1, code & 3
            if all(commutes(e, c, t) for e in out[:k]):
                cv = (code >> 6) & 15
                front[k] = np.where((regs16 & cm) == cv, regs16 ^ (1 << t), regs16)
        for k, e in deferred.items():
            for j in range(k + 1, len(out)):
                tj, cj, dj = out[j]
                if tj in (SUB_DIAG, SUB_DIAGV):
                    out[j] = [tj, cj, dj[:, e]]     # D'(r) = D(E(r))
        for k, e in front.items():
            for j in range(k):
                tj, cj, dj = out[j]
                if tj in (SUB_DIAG, SUB_DIAGV):
                    out[j] = [tj, cj, dj[:, e]]     # D'(r) = D(E(r)), E = E^-1
            load_inv = load_inv[e]
        out = [x for k, x in enumerate(out) if k not in deferred and k not in front]
        deferred = list(deferred.values())

    def flip_of(code, data):
        """(register bit, [(selector code, value)]) of a CX entry that can
        be a lane flip, else None."""
        if not flips or not data or (code >> 2) & 15:
            return None
        m = data & 3
        if m > MAX_FLIP_SEL:
            return None
        return (code & 3, [((data >> (2 + 7 * j)) & 127, (data >> (23 + j)) & 1)
                           for j in range(m)])

    def emits(k):                           # does entry k give a record
        typ, code, data = out[k]
        if typ in (SUB_DIAG, SUB_DIAGV):
            return not np.all(data == 1.0)
        return typ != SUB_CX or flip_of(code, data) is None

    def reads(k, t):                        # does entry k's record read register bit t
        typ, code, data = out[k]
        if typ in (SUB_DIAG, SUB_DIAGV):
            return (deps(data) >> t) & 1
        if typ == SUB_CX:                   # its target commutes with a flip
            return (code >> (2 + t)) & 1
        return code == t

    # Which exchanges become lane flips, and the record each rides on: the
    # first free record after it, passing only records that do not read its
    # bit; failing that, the nearest free record before it that does not
    # read its bit either (the flip commutes with every record it passes).
    # The others stay exchange records.
    carrier, virtual = {}, set()            # carrier: record -> flip
    for k, (typ, code, data) in enumerate(out):
        f = flip_of(code, data) if typ == SUB_CX else None
        if f is None:
            continue
        j = None
        for i in range(k + 1, len(out)):
            if not emits(i):
                continue
            if i not in carrier:
                j = i
                break
            if reads(i, f[0]):
                break
        if j is None:
            for i in range(k - 1, -1, -1):
                if not emits(i):
                    continue
                if reads(i, f[0]):
                    break
                if i not in carrier:
                    j = i
                    break
        if j is not None:
            carrier[j] = f
            virtual.add(k)

    def flip_word(f):
        t, sels = f
        w = 1 | (t << 1) | (sels[0][0] << 3) | (sels[0][1] << 10)
        if len(sels) > 1:
            w |= (1 << 11) | (sels[1][0] << 12) | (sels[1][1] << 19)
        return w

    nsub = 0
    active = []                             # flips applied so far
    for k, (typ, code, data) in enumerate(out):
        if k in virtual:
            continue
        if k in carrier:                    # applied before this record
            active.append(carrier[k])
        flipped = 0
        for t, _ in active:
            flipped |= 1 << t
        if typ == SUB_D1:
            ms = data[1] if isinstance(data, tuple) else [data]
            kinds_ = {d1_variant(x) for x in ms}
            typ = kinds_.pop() if len(kinds_) == 1 else SUB_D1
        if typ in (SUB_DIAG, SUB_DIAGV):
            # the flip-mask bits a table reads: slot r takes the entry of
            # slot r ^ (flips & fmask) on each lane
            sel, tab = code, data
            if np.all(tab == 1.0):
                continue
            fmask = flipped & deps(tab)
            dep = [j for j in range(REG_BITS) if (deps(tab) >> j) & 1]
            if not sel and not fmask:
                mask = 0
                for r in range(2 ** REG_BITS):
                    if tab[0, r] != 1.0:
                        mask |= 1 << r
                rec = [SUB_DIAG, 0, len(coefs), mask]
                coefs.extend(tab[0])
            else:
                cd = len(sel)
                for j, b in enumerate(sel):
                    cd |= b << (2 + 7 * j)
                # fewer than 4 bits: compact tables, entry = the slot's bits
                # under dep, in order
                if len(dep) < REG_BITS:
                    slots = [sum(((i >> a) & 1) << j for a, j in enumerate(dep))
                             for i in range(2 ** len(dep))]
                    rec = [SUB_DIAGV, cd, len(coefs),
                           16 | sum(1 << j for j in dep) | fmask << 8]
                    coefs.extend(tab[:, slots].reshape(-1))
                else:
                    rec = [SUB_DIAGV, cd, len(coefs), fmask << 8]
                    coefs.extend(tab.reshape(-1))
        elif typ == SUB_CX:
            cm = (code >> 2) & 15
            rec = [typ, code | (CX_FLIP if flipped & cm else 0), 0, data or 0]
        else:
            # one matrix, or one per selector value (a folded diagonal)
            sel, mats = data if isinstance(data, tuple) else ([], [data])
            mats = [np.asarray(x, dtype=np.complex128) for x in mats]
            word = len(sel)
            for jj, b in enumerate(sel):
                word |= b << (2 + 7 * jj)
            rec = [typ, code, len(coefs), word]
            flag = (flipped >> code) & 1
            if flag:                        # X m X for lanes holding the pair reversed
                rec[1] |= D1_FLIP
            for x in mats:
                coefs.extend(x.reshape(-1))
                if flag:
                    coefs.extend(x[::-1, ::-1].reshape(-1))
        if k in carrier:
            rec[0] |= flip_word(carrier[k]) << 8
        subs.extend(rec)
        nsub += 1
    o.coef = nsub
    store = list(range(2 ** REG_BITS))
    for e in deferred:                      # slot r stores where E(r) would
        store = [int(e[x]) for x in store]
    o.active = _affine_code([phi(x) for x in store])                  # store map
    o.phase = _affine_code(list(load_inv))                            # load map
    return o


def _gate_perm(kind, p):
    """Logical-index permutation (table over the 4 phase bits) of a monomial
    gate whose qubit j is logical bit p[j]: out index = table[in index]."""
    k = len(p)
    table = []
    for x in range(2 ** REG_BITS):
        gi = 0
        for j in range(k):
            gi |= ((x >> p[j]) & 1) << (k - 1 - j)
        # the gate maps input row src[r] to output row r
        r = int(np.nonzero(kind.src == gi)[0][0])
        y = x
        for j in range(k):
            y = (y & ~(1 << p[j])) | (((r >> (k - 1 - j)) & 1) << p[j])
        table.append(y)
    return np.asarray(table)


def _affine_code(table):
    """Encode an affine map of 4-bit indices (given as a table) as
    matrix | offset << 16: column j (bits 4j..4j+3) = image of bit j xor the
    image of 0.  Raises if the table is not affine and invertible."""
    c = int(table[0])
    cols = [int(table[1 << j]) ^ c for j in range(REG_BITS)]
    for x in range(2 ** REG_BITS):
        y = c
        for j in range(REG_BITS):
            if (x >> j) & 1:
                y ^= cols[j]
        if y != table[x]:
            raise ValueError('register map is not affine')
    if len(set(int(v) for v in table)) != 2 ** REG_BITS:
        raise ValueError('register map is not invertible')
    return sum(col << (4 * j) for j, col in enumerate(cols)) | (c << 16)


def relabel_swaps(ops, pos, low_bits):
    """Absorb a pass's uncontrolled SWAP-type gates (SWAP, ISWAP, PSWAP) by
    relabelling which tile bit holds which qubit: later ops are rewritten to
    the new bits, the gate's phases stay as a diagonal, and the pass stores
    tile bit j to index bit store_pos[j].  Swaps touching the lowest
    ``low_bits`` index bits are kept (relabelling them would break the
    coalesced runs of the store).  Returns (ops, store_pos or None)."""
    where = {b: b for b in pos}             # qubit's index bit -> tile bit holding it
    out = []
    moved = False
    for op in ops:
        bits = [where[b] for b in op.bits]
        pat = exchange_pattern(op.kind) if op.matrix is None else None
        if (pat is not None and pat[0] == 'swap' and not pat[3]
                and min(bits[pat[1]], bits[pat[2]]) >= low_bits):
            x, y = op.bits[pat[1]], op.bits[pat[2]]
            where[x], where[y] = where[y], where[x]
            moved = True
            if op.kind.phase is not None:   # phases on output rows, at the new bits
                out.append(Op([where[b] for b in op.bits],
                              bk.classify(np.diag(op.kind.phase))))
            continue
        out.append(Op(bits, op.kind, op.matrix, op.ngates))
    if not moved:
        return out, None
    holder = {t: b for b, t in where.items()}  # tile bit -> qubit index bit it holds
    return out, [holder[b] for b in pos]


def plan_blocks(ops, width):
    """Group a pass's ops (in order) into dense blocks of <= width qubits,
    with the commuting-overtake rule and bit-set growth of plan_passes.
    Returns [(bits, [ops])], bits in first-use order."""
    bits_list = [tuple(op.bits) for op in ops]
    pending = list(range(len(ops)))
    blocks = []
    while pending:
        _, taken = _grow(bits_list, pending, set(), width)
        bits = []
        for i in taken:
            bits += [b for b in bits_list[i] if b not in bits]
        blocks.append((bits, [ops[i] for i in taken]))
        done = set(taken)
        pending = [i for i in pending if i not in done]
    return blocks


def block_matrix(bits, members):
    """Dense matrix of the ops applied in order on block qubits ``bits``
    (block qubit 0 = most significant), Circuit.asgate semantics."""
    nq = len(bits)
    m = np.eye(2 ** nq, dtype=np.complex128)
    for op in members:
        m = embed(op.dense_matrix(), [bits.index(b) for b in op.bits], nq) @ m
    return m


def pack_mfma(matrix):
    """Real form R = [[Re M, -Im M], [Im M, Re M]] of a 2^k x 2^k matrix,
    packed in the operand order of the matrix-core blocks:
    P[(rb * NK + ks) * 64 + l] = R[16 rb + l % 16][4 ks + l // 16]."""
    m = np.asarray(matrix, dtype=np.complex128)
    d2 = 2 * m.shape[0]
    r = np.block([[m.real, -m.imag], [m.imag, m.real]])
    nr, nk = d2 // 16, d2 // 4
    lane = np.arange(64)
    rows = 16 * np.arange(nr)[:, None, None] + (lane % 16)[None, None, :]
    cols = 4 * np.arange(nk)[None, :, None] + (lane // 16)[None, None, :]
    return np.ascontiguousarray(r[rows, cols].reshape(-1))


def _encode_mfma(bits, matrix, pos_index, coefs):
    o = _native.QfxTileOp()
    k = len(bits)
    t = [pos_index[b] for b in bits]
    o.kind = OP_MFMA
    o.k = k
    st = sorted(t)
    for j in range(5):
        o.t[j] = t[j] if j < k else 0
        o.sorted[j] = st[j] if j < k else 0
    o.coef = len(coefs)
    coefs.extend(pack_mfma(matrix).view(np.complex128))
    o.src = o.active = o.phase = 0
    return o


MAX_STAGED = 768          # coefficients a pass without wide ops stages in LDS
_COEF_SUBS = (SUB_D1, SUB_DIAG, SUB_D1R, SUB_D1X, SUB_DIAGV)


def _op_coef_start(o, srcs):
    """First coefficient index an encoded op reads (None if it reads none)."""
    if o.kind == OP_RPHASE:
        idx = [srcs[o.src + 4 * q + 2] for q in range(o.coef)
               if srcs[o.src + 4 * q] & REC_TYPE in _COEF_SUBS]
        return min(idx) if idx else None
    return o.coef


def _rebase_coef(o, srcs, off):
    if o.kind == OP_RPHASE:
        for q in range(o.coef):
            if srcs[o.src + 4 * q] & REC_TYPE in _COEF_SUBS:
                srcs[o.src + 4 * q + 2] -= off
    else:
        o.coef -= off


_SWAP_KIND = None


def _swap_kind():
    global _SWAP_KIND
    if _SWAP_KIND is None:
        _SWAP_KIND = bk.classify(np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]],
                                          dtype=np.complex128))
    return _SWAP_KIND


def rename_swaps(kinds, bits_list, n):
    """Remove uncontrolled SWAP-type gates (SWAP, ISWAP, PSWAP) from a gate
    list by renaming: after a swap of index bits x and y the data of x sits
    where y's was, so later gates act on the swapped bits and the gate moves
    nothing.  Its phases (ISWAP, PSWAP) stay as a diagonal on the renamed
    bits.  Returns (kinds, bits, where, number removed): where[b] is the
    index bit holding b's data after the last gate."""
    where = list(range(n))
    out_k, out_b = [], []
    removed = 0
    for kd, bits in zip(kinds, bits_list):
        pat = exchange_pattern(kd) if kd.kind == bk.GateKind.MONOMIAL else None
        if pat is not None and pat[0] == 'swap' and not pat[3]:
            x, y = bits[pat[1]], bits[pat[2]]
            where[x], where[y] = where[y], where[x]
            removed += 1
            if kd.phase is not None:        # phases on output rows, at the new bits
                out_k.append(bk.classify(np.diag(kd.phase)))
                out_b.append(tuple(where[b] for b in bits))
            continue
        out_k.append(kd)
        out_b.append(tuple(where[b] for b in bits))
    return out_k, out_b, where, removed


def restoring_swaps(where, pivot='low'):
    """Plain swaps (as index-bit pairs, in order) that move the data of every
    bit b from where[b] back to b.  Each walks one cycle of the
    displacement: swapping the data at where[b] into b fixes b and passes
    the displaced data along the cycle, so every swap of the cycle shares
    the position its walk starts from, one swap per bit of the cycle but
    one.  pivot 'low' starts where the cycle's lowest position holds its
    data, making that shared position a low bit when the cycle has one (the
    low bits are in every tile, so each swap needs one more tile bit only);
    'high' starts at its highest position, keeping the swaps off the low
    bits (pass-level store permutations instead of register relabelling)."""
    where = list(where)
    holder = {p: b for b, p in enumerate(where)}     # position -> bit whose data is there
    out = []
    for b0 in range(len(where)):
        if where[b0] == b0:
            continue
        cyc, b = [b0], holder[b0]
        while b != b0:
            cyc.append(b)
            b = holder[b]
        b = (min if pivot == 'low' else max)(cyc, key=lambda c: where[c])
        while where[b] != b:
            p = where[b]
            out.append((p, b))
            other = holder[b]               # the bit whose data sat at b moves to p
            where[other], holder[p] = p, other
            where[b], holder[b] = b, b
            b = other
    return out


def virtualize_swaps(kinds, bits_list, n, pivot='low'):
    """rename_swaps, then the restoring swaps appended: the same circuit.
    Returns (kinds, bits, number of swaps removed)."""
    out_k, out_b, where, removed = rename_swaps(kinds, bits_list, n)
    for pq in restoring_swaps(where, pivot):
        out_k.append(_swap_kind())
        out_b.append(pq)
    return out_k, out_b, removed


def plan_passes_restoring(kinds, bits_list, where, n, T, L, prepare, pivot='low'):
    """plan_passes (commutation-aware) for a gate list renamed by
    rename_swaps, whose data ends displaced (where[b]: the index bit holding
    b's data after the last gate).  After each pass is chosen, its store
    permutation also moves data home where that costs nothing: for tile
    positions p above the low L bits whose data belongs (at the end) to a
    bit q that is also such a tile position, the pass stores p to q (plain
    swaps appended to the pass, which relabel_swaps absorbs), and every
    later gate is renamed accordingly.  The swaps that restore what is
    still displaced trail the pending gates, recomputed after each pass,
    and join passes like any gate.  ``prepare(kinds, bits)`` gives
    (tileable, free, need, types) per gate.  Returns (passes, kinds, bits):
    the kinds and bits lists extended by the swaps the passes hold."""
    kinds, bits_list = list(kinds), list(bits_list)
    tileable, free, need, types = (list(x) if x is not None else None
                                   for x in prepare(kinds, bits_list))
    low = set(range(min(L, n)))
    final = list(where)                     # bit -> index bit holding its data at the end

    def add_gate(kd, bits):
        t, f, nd, ty = prepare([kd], [bits])
        kinds.append(kd)
        bits_list.append(tuple(bits))
        tileable.append(t[0])
        if free is not None:
            free.append(f[0])
        if need is not None:
            need.append(nd[0] if nd is not None else tuple(bits))
        types.append(ty[0])
        return len(kinds) - 1

    def rename(i, m):
        bits = tuple(m.get(b, b) for b in bits_list[i])
        bits_list[i] = bits
        t, f, nd, ty = prepare([kinds[i]], [bits])
        if need is not None:
            need[i] = nd[0] if nd is not None else bits
        types[i] = ty[0]

    def tail_swaps():
        return [add_gate(_swap_kind(), pq) for pq in restoring_swaps(final, pivot)]

    tail = set(tail_swaps())
    pending = list(range(len(kinds) - len(tail))) + sorted(tail)
    passes = []
    while pending:
        first = pending[0]
        if not tileable[first]:
            passes.append((None, [first]))
            pending = pending[1:]
            continue
        S, _ = _grow(bits_list, pending, set(low), T, tileable, free=free, types=types,
                     need=need)
        b = 0
        while len(S) < T:
            if b not in S:
                S.add(b)
            b += 1
        taken = _absorb(bits_list, pending, S, tileable, free, types, need)
        done = set(taken)
        # restoring swaps taken by this pass have moved their data
        for i in taken:
            if i in tail:
                p, q = bits_list[i]
                final = [q if f == p else p if f == q else f for f in final]
        # store permutation: move data home among the tile's positions above
        # its contiguous low run (those keep the store coalesced)
        run = 0
        while run in S:
            run += 1
        upper = {q for q in S if q >= max(run, L)}
        holder = {f: bb for bb, f in enumerate(final)}
        sigma = {}                          # data at position p goes to sigma[p]
        for pos in sorted(upper):
            dest = holder[pos]              # the bit whose data ends at pos
            if dest != pos and dest in upper:
                sigma[pos] = dest
        extra = []
        if sigma:
            # complete the partial map to a permutation of the upper positions
            free_dst = sorted(upper - set(sigma.values()))
            for pos in sorted(upper - set(sigma)):
                d = free_dst.pop(0)
                if d != pos:
                    sigma[pos] = d
            # each cycle c0 -> c1 -> ... as swaps (c0, c1), (c0, c2), ...
            seen = set()
            for c0 in sorted(sigma):
                if c0 in seen:
                    continue
                c = sigma[c0]
                seen.add(c0)
                while c != c0:
                    extra.append(add_gate(_swap_kind(), (c0, c)))
                    seen.add(c)
                    c = sigma[c]
            final = [sigma.get(f, f) for f in final]
        passes.append((sorted(S), taken + extra))
        pending = [i for i in pending if i not in done and i not in tail]
        for i in pending:
            if sigma:
                rename(i, sigma)
        tail = set(tail_swaps())
        pending += sorted(tail)
    return passes, kinds, bits_list


class HostPlan:
    """A gate list compiled into tiled passes on the host: the pass list,
    the op records, coefficient and source tables, validated by the
    library's qfx_tile_check_ops.  ``passes`` holds ('tile', (tile bits,
    first op, op count, block, flags, first coefficient, coefficient count))
    and ('gate', (bits, GateKind)) for gates too wide to tile; op records
    index their pass's coefficient table."""

    def __init__(self, gates, qubits, tile_bits=MAX_TILE_BITS, low_bits=3, block=256,
                 fuse_width=4, mode='reg', persistent=True, reg_io=True, reg_in=True,
                 free_controls=True, virtual_swaps=True, lane_flips=True, store_cx=True,
                 merge_records=True):
        qubits = tuple(qubits)
        n = len(qubits)
        self.n = n
        self.qubits = qubits
        self.tile_bits = T = min(tile_bits, n)
        self.low_bits = low_bits
        from .autotune import block_for
        self.block = block_for(T, block)
        self.mode = mode if T >= REG_BITS + 2 else 'lds'
        if mode == 'mfma' and T < max(fuse_width, MFMA_MIN_BITS) + 4:
            self.mode = 'lds'           # tile too small for 16-group columns
        # register phases and matrix-core blocks run in blocks of <= 512
        # lanes; 4- and 5-qubit ops only in blocks of <= 256
        if self.mode in ('reg', 'mfma') and self.block > 512:
            self.block = 512
        self.fuse_width = fuse_width
        self.persistent = persistent
        self.reg_io = reg_io
        self.reg_in = reg_in
        self.free_controls = free_controls
        self.lane_flips = lane_flips
        self.store_cx = store_cx
        self.merge_records = merge_records
        self.gates = list(gates)
        bits0 = [tuple(n - 1 - qubits.index(q) for q in g.qubits) for g in gates]
        kinds0 = [g.kind for g in gates]

        def prepare(kinds, bits_list):
            """tileable, free, need and commutation types of a gate list"""
            tileable = [len(b) <= 5 and len(b) <= T for b in bits_list]
            # register phases apply diagonal gates whatever the tile bits:
            # the phases of the qubits outside the tile are selected per tile
            # (only where no pass can turn wide, i.e. every gate acts on <= 3
            # qubits), and controlled X gates need only their target in the
            # tile: controls outside it are read from the tile's index per tile
            free = need = None
            if self.mode == 'reg' and all(len(b) <= 3 for b in bits_list):
                free = [kd.kind == bk.GateKind.DIAGONAL for kd in kinds]
                if free_controls:
                    need = []
                    for kd, b in zip(kinds, bits_list):
                        cx = cx_parts(kd, b)
                        need.append((cx[0],) if cx else b)
            types = [gate_types(kd, b) for kd, b in zip(kinds, bits_list)]
            return tileable, free, need, types

        self.n_virtual = 0
        if self.mode == 'reg' and virtual_swaps:
            # uncontrolled SWAP-type gates move no data where they stand:
            # later gates are renamed onto the swapped bits, and swaps
            # appended at the end restore the order.  Both walks of the
            # restoring cycles are planned (commutation-aware); the fewer
            # passes win, ties going to the high walk (fewer phases)
            rk, rb, where, self.n_virtual = rename_swaps(kinds0, bits0, n)
            best = None
            for pivot in ('high', 'low'):
                planned, kinds, bits_list = plan_passes_restoring(
                    rk, rb, where, n, T, low_bits, prepare, pivot=pivot)
                if best is None or len(planned) < len(best[0]):
                    best = (planned, kinds, bits_list)
            planned, kinds, bits_list = best
            tileable, free, need, types = prepare(kinds, bits_list)
        else:
            kinds, bits_list = kinds0, bits0
            tileable, free, need, types = prepare(kinds, bits_list)
            # commutation-aware overtaking can pack the same gates into fewer
            # passes, or into as many heavier ones: keep it only when it saves
            # a pass
            planned = plan_passes(bits_list, n, T, low_bits, tileable, free=free, need=need)
            typed = plan_passes(bits_list, n, T, low_bits, tileable, free=free, types=types,
                                need=need)
            if len(typed) < len(planned):
                planned = typed
        self.passes = []           # (kind, payload)
        ops_host, coefs, srcs = [], [], []
        self.n_phases = 0
        self.n_blocks = 0
        orders, holders = [], []   # per tile pass: tile bits, qubit each holds at the end
        for pos, members in planned:
            if pos is None:
                self.passes.append(('gate', (list(bits_list[members[0]]), kinds[members[0]])))
                continue
            tset = set(pos)
            pos = sorted(pos)
            pos_index = {b: j for j, b in enumerate(pos)}
            ops = []
            for i in members:
                bits = bits_list[i]
                if tset.issuperset(bits):
                    ops.append(Op(bits, kinds[i]))
                elif kinds[i].kind == bk.GateKind.DIAGONAL:
                    sd = SelDiagonal(kinds[i], bits, tset)
                    ops.append(Op(sd.in_bits, sd))
                else:                       # controlled X, controls outside the tile
                    t, ctl = cx_parts(kinds[i], bits)
                    sc = SelCX(t, [c for c in ctl if c in tset], [c for c in ctl if c not in tset])
                    ops.append(Op(sc.in_bits, sc))
            wide = any(len(op.bits) > 3 for op in ops)
            start = len(ops_host)
            c_lo = len(coefs)
            store_pos = None
            if self.mode == 'mfma':
                width = max(MFMA_MIN_BITS, min(fuse_width, 5))
                for bits, group in plan_blocks(ops, width):
                    m = block_matrix(bits, group)
                    # pad narrow blocks with free tile bits (identity factors)
                    pad = [b for b in pos if b not in bits][:max(0, MFMA_MIN_BITS - len(bits))]
                    if pad:
                        m = np.kron(m, np.eye(2 ** len(pad)))
                        bits = bits + pad
                    ops_host.append(_encode_mfma(bits, m, pos_index, coefs))
                    self.n_blocks += 1
                block, flags = self.block, TILE_MFMA
            elif self.mode == 'reg' and not wide:
                ops, store_pos = relabel_swaps(ops, pos, low_bits)
                # diagonals and controls free of the register bits; a phase
                # whose tables exceed the staging limit runs as consecutive
                # phases on the same bits, each holding part of its gates
                for item in plan_phases(ops, free_diagonals=free is not None):
                    if item[0] == 'phase':
                        for o in self._encode_fitting(item[1], item[2], pos_index, T,
                                                      coefs, srcs, lane_flips, store_cx,
                                                      merge_records):
                            ops_host.append(o)
                            self.n_phases += 1
                    else:
                        ops_host.append(_encode_op(item[1], pos_index, coefs, srcs))
                block, flags = self.block, 0
                # a last register phase stores straight to HBM when its lanes
                # run over the low (contiguous) tile bits
                last = ops_host[-1] if len(ops_host) > start else None
                if (last is not None and persistent and self.reg_io and last.kind == OP_RPHASE
                        and (1 << T) == 16 * block and block <= 512
                        and min(last.t[j] for j in range(REG_BITS)) >= min(low_bits, 3)):
                    flags = TILE_REG_IO
                # likewise a first register phase loads straight from HBM
                first = ops_host[start] if len(ops_host) > start else None
                if (first is not None and persistent and self.reg_in and first.kind == OP_RPHASE
                        and (1 << T) == 16 * block and block <= 512
                        and min(first.t[j] for j in range(REG_BITS)) >= min(low_bits, 3)):
                    flags |= TILE_REG_IN
            else:
                block = min(self.block, 256) if wide else self.block
                width = self.fuse_width if block <= 256 else min(self.fuse_width, 3)
                for op in fuse_ops(ops, width):
                    ops_host.append(_encode_op(op, pos_index, coefs, srcs))
                wide = wide or any(o.k > 3 for o in ops_host[start:] if o.kind != OP_RPHASE)
                flags = TILE_WIDE if wide else 0
                if wide:
                    block = min(block, 256)
            if persistent:
                flags |= TILE_PERSISTENT
            orders.append(pos)
            holders.append({b: (store_pos[j] if store_pos is not None else b)
                            for j, b in enumerate(pos)})
            self._add_tile_passes(len(orders) - 1, ops_host, start, coefs, c_lo, srcs, block,
                                  flags)
        # each tile pass's store permutation: the last launch of a pass
        # stores every tile bit to the index bit of the qubit it then holds
        for i, (kind, payload) in enumerate(self.passes):
            if kind != 'tile':
                continue
            k, start, nops, block, flags, c_lo, ncoef, last = payload
            hold = holders[k]
            dest = None
            if last and any(hold[b] != b for b in orders[k]):
                dest = list(range(n))
                for b in orders[k]:
                    dest[b] = hold[b]
            self.passes[i] = ('tile', (orders[k], start, nops, block, flags, c_lo, ncoef, dest))
        self.n_passes = len(self.passes)
        self.n_ops = len(ops_host)
        srcs.extend([0, 0, 0, 0])       # the sub-op loop reads one record ahead,
        # also at the src of a phase without sub-ops
        self.ops_host, self.coefs, self.srcs = ops_host, coefs, srcs
        prepare_ops(ops_host)
        # validate every pass with its own block and flags
        lib = _native.load()
        src_arr = _native.int_array(srcs)
        for kind, payload in self.passes:
            if kind != 'tile':
                continue
            _, start, nops, block, flags, c_lo, ncoef, _ = payload
            arr = (_native.QfxTileOp * max(1, nops))(*ops_host[start:start + nops])
            _native.check(lib.qfx_tile_check_ops(T, block, flags, arr, nops, ncoef,
                                                 src_arr, len(srcs)), 'qfx_tile_check_ops')

    @staticmethod
    def _encode_fitting(R, members, pos_index, T, coefs, srcs, flips=True, store_cx=True,
                        merge_records=True):
        """Register phase(s) for ``members`` on bits R, each staging at most
        MAX_STAGED coefficients: a phase over the limit is cut into two
        consecutive phases of its first and second half (each leaves the
        tile in logical order, so they compose)."""
        c0, s0 = len(coefs), len(srcs)
        o = _encode_phase(R, members, pos_index, T, coefs, srcs, flips, store_cx, merge_records)
        if len(coefs) - c0 <= MAX_STAGED or len(members) == 1:
            return [o]
        del coefs[c0:], srcs[s0:]
        h = len(members) // 2
        return (HostPlan._encode_fitting(R, members[:h], pos_index, T, coefs, srcs, flips,
                                         store_cx, merge_records)
                + HostPlan._encode_fitting(R, members[h:], pos_index, T, coefs, srcs, flips,
                                           store_cx, merge_records))

    def _add_tile_passes(self, k, ops_host, start, coefs, c_lo, srcs, block, flags):
        """Record tile pass k (ops_host[start:]) with coefficient indices
        made relative to the pass's table; a pass that would stage more than
        MAX_STAGED coefficients is cut into consecutive launches on the same
        tile bits, of which only the last stores to the next bit order (the
        payload's last field).  The load positions and store permutations are
        filled in once every pass's bit order is known."""
        ops = ops_host[start:]
        if not ops:
            # only absorbed exchanges (or a change of bit order): the pass is
            # its store permutation
            self.passes.append(('tile', (k, start, 0, block,
                                         flags & ~(TILE_REG_IO | TILE_REG_IN), c_lo,
                                         0, True)))
            return
        bounds = [_op_coef_start(o, srcs) for o in ops] + [len(coefs)]
        # an op without coefficients starts where the next op starts
        for i in range(len(ops) - 1, -1, -1):
            if bounds[i] is None:
                bounds[i] = bounds[i + 1]
        staged = not flags & (TILE_WIDE | TILE_MFMA)
        a = 0
        while a < len(ops):
            b = a + 1 if staged else len(ops)
            while b < len(ops) and bounds[b + 1] - bounds[a] <= MAX_STAGED:
                b += 1
            for o in ops[a:b]:
                _rebase_coef(o, srcs, bounds[a])
            fl = flags
            if fl & TILE_REG_IO and ops[b - 1].kind != OP_RPHASE:
                fl &= ~TILE_REG_IO
            if fl & TILE_REG_IN and (a > 0 or ops[a].kind != OP_RPHASE):
                fl &= ~TILE_REG_IN
            self.passes.append(('tile', (k, start + a, b - a, block, fl, bounds[a],
                                         bounds[b] - bounds[a], b == len(ops))))
            a = b

    # real FP64 operations per amplitude: a complex multiply is 6, a complex
    # multiply-add 8; a 2x2 real matrix on a complex pair is 6 per output
    _SUB_FLOPS = {SUB_D1: 14, SUB_D1R: 6, SUB_D1X: 6, SUB_DIAG: 6, SUB_DIAGV: 6, SUB_CX: 0}

    def flops(self):
        """Algorithmic FP64 operations of the plan's tile passes, per pass
        launch in order: 2^n times the per-amplitude count of every op (a
        register phase: its sub-ops; a dense k-qubit op: 2^k complex
        multiply-adds less one add, 8 * 2^k - 2; a sparse op: one complex
        multiply; a matrix-core block: 8 * 2^k)."""
        out = []
        for kind, payload in self.passes:
            if kind != 'tile':
                continue
            _, start, nops, *_ = payload
            per_amp = 0
            for o in self.ops_host[start:start + nops]:
                if o.kind == OP_RPHASE:
                    per_amp += sum(self._SUB_FLOPS[self.srcs[o.src + 4 * q] & REC_TYPE]
                                   for q in range(o.coef))
                elif o.kind == OP_DENSE:
                    per_amp += 8 * 2 ** o.k - 2
                elif o.kind == OP_SPARSE:
                    per_amp += 6
                else:
                    per_amp += 8 * 2 ** o.k
            out.append(per_amp * 2 ** self.n)
        return out

    def stats(self):
        ng = len(self.gates)
        return {'gates': ng, 'passes': self.n_passes, 'ops': self.n_ops,
                'phases': self.n_phases, 'blocks': self.n_blocks, 'mode': self.mode,
                'persistent': self.persistent,
                'tile_bits': self.tile_bits, 'low_bits': self.low_bits,
                'block': self.block, 'fuse_width': self.fuse_width}


def prepare_ops(ops):
    """Fill each register phase's derived LDS words in place
    (qfx_tile_prepare_ops); run after the last change to the ops."""
    lib = _native.load()
    for o in ops:
        _native.check(lib.qfx_tile_prepare_ops(_native.ctypes.byref(o), 1),
                      'qfx_tile_prepare_ops')


class TilePlan(HostPlan):
    """A HostPlan uploaded to the device once; running it is one kernel
    launch per pass."""

    def __init__(self, gates, qubits, **cfg):
        super().__init__(gates, qubits, **cfg)
        ops_host, coefs, srcs = self.ops_host, self.coefs, self.srcs
        arr = (_native.QfxTileOp * max(1, self.n_ops))(*ops_host)
        raw = bytes(arr)
        dev = bk.DEVICE
        self._ops = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(dev)
        self._coef = torch.from_numpy(np.asarray(coefs if coefs else [0j],
                                                 dtype=np.complex128)).to(dev)
        self._srcs = torch.tensor(srcs if srcs else [0], dtype=torch.int32, device=dev)
        self._opsize = _native.ctypes.sizeof(_native.QfxTileOp)
        self._pos_arrays = [(_native.int_array(p[1][0]), len(p[1][0]),
                             _native.int_array(p[1][7]) if p[1][7] is not None else None)
                            if p[0] == 'tile' else None for p in self.passes]

    def run_inplace(self, work, prof=None):
        """Run every pass in place on the flat complex128 device state."""
        s = bk.stream_handle()
        n = self.n
        lib = _native.load()
        wp = _native.ctypes.c_void_p(work.data_ptr())
        base_ops = self._ops.data_ptr()
        coef = self._coef.data_ptr()
        srcs = _native.ctypes.c_void_p(self._srcs.data_ptr())
        for (kind, payload), posarr in zip(self.passes, self._pos_arrays):
            if kind == 'gate':
                bits, gk = payload
                tok = prof.start(kernel_class(gk)) if prof is not None else None
                bk.apply_gate(work, work, n, bits, gk, inplace_sparse=True)
            else:
                pos, start, nops, block, flags, c_lo, ncoef, _ = payload
                tok = prof.start('tile_pass') if prof is not None else None
                rc = lib.qfx_tile_pass(s, wp, n, posarr[1], posarr[0], posarr[2],
                                       _native.ctypes.c_void_p(base_ops + start * self._opsize),
                                       nops, _native.ctypes.c_void_p(coef + 16 * c_lo), ncoef,
                                       srcs, block, flags)
                _native.check(rc, 'qfx_tile_pass')
            if tok is not None:
                prof.stop(tok)
        return work

    def run(self, ket: State) -> State:
        if tuple(ket.qubits) != self.qubits:
            raise ValueError('plan was built for qubits %s' % (self.qubits,))
        work = bk.as_device(ket.tensor).clone()
        self.run_inplace(work)
        return State(work.view([2] * self.n), self.qubits, ket.memory)


# Plans of recent gate lists.  A plan depends only on its gates (their
# matrices are fixed when a gate is built), qubits and configuration; the
# entry keeps its gates alive, so the ids in its key cannot be reused by
# other objects while it is cached.
_PLAN_CACHE = OrderedDict()
_PLAN_CACHE_SIZE = 8
_plan_lock = threading.Lock()


def plan_for(gates, qubits, **kwargs):
    """Fused execution plan for ``gates`` on a state with these qubits, or
    None when the state is too small to tile (fewer than 6 qubits).  The
    last few plans are kept, so running the same gate objects again (a
    circuit run repeatedly) skips the planner."""
    n = len(qubits)
    if n < MIN_TILE_BITS:
        return None
    from . import autotune
    cfg = autotune.choose(n, len(gates))
    cfg.update(kwargs)
    gates = tuple(gates)
    key = (tuple(id(g) for g in gates), tuple(qubits), tuple(sorted(cfg.items())),
           str(bk.DEVICE))
    with _plan_lock:
        hit = _PLAN_CACHE.get(key)
        if hit is not None:
            _PLAN_CACHE.move_to_end(key)
            return hit[1]
    plan = TilePlan(gates, qubits, **cfg)
    with _plan_lock:
        _PLAN_CACHE[key] = (gates, plan)
        while len(_PLAN_CACHE) > _PLAN_CACHE_SIZE:
            _PLAN_CACHE.popitem(last=False)
    return plan

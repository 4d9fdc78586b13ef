"""Host-side checks of the fusion planner (no GPU): the passes it builds,
replayed op by op with the oracle, give the circuit's state; every op fits
its tile; merged blocks equal the product of their gates; the encoded op
tables pass the library's own validator."""
import numpy as np
import pytest

from oracle import qf_oracle as O
from quantumflow_amd import fusion, _native
from quantumflow_amd import backend as bk
from quantumflow_amd.workloads import random_mixed_circuit, layered_rotation_circuit, build_gates


def replay(gates, n, T, L, width):
    """Apply the planned (and fused) ops with the oracle, in plan order."""
    qubits = tuple(range(n))
    bits_list = [tuple(n - 1 - q for q in g.qubits) for g in gates]
    passes = fusion.plan_passes(bits_list, n, T, L)
    ket = O.zero_state(n)
    covered = []
    for pos, members in passes:
        assert pos is not None and len(pos) == T and pos == sorted(pos)
        assert set(range(min(L, n))) <= set(pos)
        ops = [fusion.Op(bits_list[i], gates[i].kind) for i in members]
        for op in ops:
            assert set(op.bits) <= set(pos)
        fused = fusion.fuse_ops(ops, width)
        assert sum(o.ngates for o in fused) == len(members)
        for op in fused:
            assert len(op.bits) <= max(width, 3)
            axes = [n - 1 - b for b in op.bits]
            m = op.dense_matrix()
            k = len(axes)
            ket = O.tensormul(np.asarray(m).reshape([2] * (2 * k)), ket, axes)
        covered += members
    assert sorted(covered) == list(range(len(gates)))
    return ket


@pytest.mark.parametrize("T,L,width", [(6, 3, 1), (7, 2, 3), (8, 3, 4), (9, 4, 5)])
def test_plan_replay_matches_circuit(T, L, width):
    n = 10
    desc = random_mixed_circuit(n, 12, seed=T)
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    got = replay(gates, n, T, L, width)
    assert np.max(np.abs(got - ref)) < 1e-12


def test_plan_replay_layered():
    n = 12
    desc = layered_rotation_circuit(n, 6, seed=4)
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    got = replay(gates, n, 12, 3, 4)
    assert np.max(np.abs(got - ref)) < 1e-12


def test_pass_count_28q():
    desc = random_mixed_circuit(28, 40, seed=2024)
    bits = [tuple(27 - q for q in g[2]) for g in desc]
    passes = fusion.plan_passes(bits, 28, 13, 3)
    assert len(passes) <= 60          # vs 728 gates unfused


def test_embed_closed_form():
    rng = np.random.default_rng(0)
    u = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    # gate on block qubits (2, 0) of 3 == full_operator on axes [2, 0]
    e = fusion.embed(u, [2, 0], 3)
    assert np.allclose(e, O.full_operator(u, 3, [2, 0]), atol=1e-15)


def test_encoded_ops_validate():
    desc = random_mixed_circuit(10, 6, seed=1)
    gates = build_gates(desc)
    n, T = 10, 8
    bits_list = [tuple(n - 1 - q for q in g.qubits) for g in gates]
    coefs, srcs, ops = [], [], []
    for pos, members in fusion.plan_passes(bits_list, n, T, 3):
        pidx = {b: j for j, b in enumerate(pos)}
        for op in fusion.fuse_ops([fusion.Op(bits_list[i], gates[i].kind) for i in members], 3):
            ops.append(fusion._encode_op(op, pidx, coefs, srcs))
    arr = (_native.QfxTileOp * len(ops))(*ops)
    lib = _native.load()
    assert lib.qfx_tile_check_ops(T, 256, 0, arr, len(ops), len(coefs), _native.int_array(srcs),
                                  len(srcs)) == 0
    # a target outside the tile is rejected
    bad = (_native.QfxTileOp * 1)(ops[0])
    bad[0].t[0] = T
    bad[0].sorted[0] = T
    assert lib.qfx_tile_check_ops(T, 256, 0, bad, 1, len(coefs), _native.int_array(srcs),
                                  len(srcs)) == _native.QFX_EINVAL


def test_empty_phase_needs_a_record():
    """A register phase without sub-ops still has its first record read
    ahead, so the validator wants one record at its src; plans append one."""
    T = 8
    pidx = {b: b for b in range(T)}
    coefs, srcs = [], []
    o = fusion._encode_phase([0, 1, 2, 3], [], pidx, T, coefs, srcs)
    assert o.coef == 0 and o.src == len(srcs)
    fusion.prepare_ops([o])
    arr = (_native.QfxTileOp * 1)(o)
    lib = _native.load()
    assert lib.qfx_tile_check_ops(T, 256, 0, arr, 1, 0, _native.int_array(srcs + [0]),
                                  len(srcs)) == _native.QFX_EINVAL
    padded = srcs + [0, 0, 0, 0]
    assert lib.qfx_tile_check_ops(T, 256, 0, arr, 1, 0, _native.int_array(padded),
                                  len(padded)) == 0
    gates = build_gates(random_mixed_circuit(10, 4, seed=3))
    plan = fusion.HostPlan(gates, tuple(range(10)), tile_bits=8, low_bits=3, block=256)
    assert plan.srcs[-4:] == [0, 0, 0, 0]


def test_gatekind_classification():
    assert bk.classify(O.gate_matrix('CNOT')).kind == 'monomial'
    assert bk.classify(O.gate_matrix('CNOT')).phase is None
    assert bk.classify(O.gate_matrix('ISWAP')).kind == 'monomial'
    assert bk.classify(O.gate_matrix('RZ', 0.3)).kind == 'diagonal'
    assert bk.classify(O.gate_matrix('H')).kind == 'dense'
    assert bk.classify(O.gate_matrix('CCNOT')).kind == 'monomial'


def unpack_mfma(slots, k):
    """Complex matrix from the packed real form, by the header's formula."""
    p = np.asarray(slots, dtype=np.complex128).view(np.float64)
    d2 = 2 << k
    nr, nk = d2 // 16, d2 // 4
    r = np.zeros((d2, d2))
    for rb in range(nr):
        for ks in range(nk):
            for lane in range(64):
                r[16 * rb + lane % 16, 4 * ks + lane // 16] = p[(rb * nk + ks) * 64 + lane]
    d = 1 << k
    assert np.array_equal(r[:d, :d], r[d:, d:]) and np.array_equal(r[:d, d:], -r[d:, :d])
    return r[:d, :d] + 1j * r[d:, :d]


def apply_store(ket, n, pos, dest):
    """A pass's store permutation: index bit b is written to index bit
    dest[b] (None: unchanged)."""
    if dest is None:
        return ket
    assert sorted(dest) == list(range(n))
    axes = list(range(n))
    for b in range(n):
        axes[n - 1 - dest[b]] = n - 1 - b
    return np.transpose(ket, axes)


def _phase_unitary(o, coefs, srcs, sel_val):
    """16x16 operator of register phase o on its register bits (register bit
    j <-> tile bit t[j]); sel_val maps an index bit outside the tile to its
    value (selector diagonals)."""
    u = np.eye(16, dtype=complex)
    d2 = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    fm = 0                                  # the lane's flip mask

    def holds(w):                           # selector j must equal bit 23 + j
        return all(sel_val[(w >> (2 + 7 * j)) & 127] == (w >> (23 + j)) & 1
                   for j in range(w & 3))

    def flip(f):                            # a lane-flip word riding on a record
        sels = [((f >> 3) & 127, (f >> 10) & 1)]
        if f & 2048:
            sels.append(((f >> 12) & 127, (f >> 19) & 1))
        return (1 << ((f >> 1) & 3)) if all(sel_val[c] == val for c, val in sels) else 0
    for q in range(o.coef):
        typ, code, c0, mask = srcs[o.src + 4 * q: o.src + 4 * q + 4]
        if typ >> 8:                        # a lane flip applies before the record
            fm ^= flip(typ >> 8)
        typ &= fusion.REC_TYPE
        if typ == fusion.SUB_DIAG:
            ph = np.where([(mask >> r) & 1 for r in range(16)], coefs[c0:c0 + 16], 1.0)
            u = np.diag(ph) @ u
            continue
        if typ == fusion.SUB_DIAGV:
            m = code & 3
            v = sum(sel_val[(code >> (2 + 7 * j)) & 127] << j for j in range(m))
            rs = np.arange(16) ^ (fm & (mask >> 8) & 15)   # slots the lane reads
            if mask & 16:                   # compact tables over register bits d
                d = [j for j in range(4) if (mask >> j) & 1]
                idx = [sum(((r >> j) & 1) << a for a, j in enumerate(d)) for r in rs]
                u = np.diag(coefs[c0 + (v << len(d)) + np.asarray(idx)]) @ u
            else:
                u = np.diag(coefs[c0 + 16 * v + rs]) @ u
            continue
        if typ == fusion.SUB_CX:
            # selector controls (mask): selector j must equal bit 23 + j
            if mask and not holds(mask):
                continue
            perm = np.zeros((16, 16))
            t, cm, cv = code & 3, (code >> 2) & 15, (code >> 6) & 15
            if code & fusion.CX_FLIP:
                cv ^= fm & cm
            for r in range(16):
                dst = r ^ (1 << t) if (r & cm) == cv else r
                perm[dst, r] = 1
            u = perm @ u
            continue
        if typ in (fusion.SUB_D1, fusion.SUB_D1R, fusion.SUB_D1X):
            bits = [code & 3]
        else:
            bits = list(d2[code])
        k = len(bits)
        if k == 1 and mask:                 # one matrix per selector value
            v = sum(sel_val[(mask >> (2 + 7 * j)) & 127] << j for j in range(mask & 3))
            c0 += v * (8 if code & fusion.D1_FLIP else 4)
        if k == 1 and code & fusion.D1_FLIP and (fm >> bits[0]) & 1:
            c0 += 4                         # X m X: the lane holds the pair reversed
        m = coefs[c0:c0 + 4 ** k].reshape(2 ** k, 2 ** k)
        # the specialised kernels read only part of each coefficient
        if typ == fusion.SUB_D1R:
            m = m.real.astype(complex)
        elif typ == fusion.SUB_D1X:
            m = np.array([[m[0, 0].real, 1j * m[0, 1].imag],
                          [1j * m[1, 0].imag, m[1, 1].real]])
        axes = [3 - b for b in reversed(bits)]      # gate MSB <-> highest bit
        u = O.full_operator(m, 4, axes) @ u

    # load map: register r starts with logical slot L(r); store map:
    # register r goes to logical slot S(r)
    def affine(m, r):
        y = m >> 16
        for j in range(4):
            if (r >> j) & 1:
                y ^= (m >> (4 * j)) & 15
        return y
    load = np.zeros((16, 16))
    store = np.zeros((16, 16))
    for r in range(16):
        load[r, affine(o.phase, r)] = 1
        # register r holds what register r ^ fm holds without flips
        store[affine(o.active, r ^ fm), r] = 1
    return store @ u @ load


def _emulate_ops(ket, n, pos, ops, coefs, srcs):
    """Apply encoded qfx_tile_op records to a numpy state, as the tile kernel
    defines them (register phases: register bit j <-> tile bit t[j];
    selector diagonals read the index bits outside the tile)."""
    coefs = np.asarray(coefs)
    for o in ops:
        if o.kind == fusion.OP_RPHASE:
            gb = [pos[o.t[j]] for j in range(fusion.REG_BITS)]
            sel = set()                     # selector codes of the phase
            for q in range(o.coef):
                typ, code = srcs[o.src + 4 * q], srcs[o.src + 4 * q + 1]
                f = typ >> 8                # a lane flip's selectors
                if f:
                    sel |= {(f >> 3) & 127} | ({(f >> 12) & 127} if f & 2048 else set())
                typ &= fusion.REC_TYPE
                if typ == fusion.SUB_DIAGV:
                    sel |= {(code >> (2 + 7 * j)) & 127 for j in range(code & 3)}
                if typ in (fusion.SUB_D1, fusion.SUB_D1R, fusion.SUB_D1X):
                    w = srcs[o.src + 4 * q + 3]
                    sel |= {(w >> (2 + 7 * j)) & 127 for j in range(w & 3)}
                if typ == fusion.SUB_CX:
                    w = srcs[o.src + 4 * q + 3]
                    sel |= {(w >> (2 + 7 * j)) & 127 for j in range(w & 3)}
            sel = sorted(sel)

            def index_bit(c):               # selector code -> state-index bit
                return c if c < 64 else pos[c - 64]
            assert not {c for c in sel if c < 64} & set(pos)
            assert not {index_bit(c) for c in sel if c >= 64} & set(gb)
            axes = [n - 1 - gb[j] for j in (3, 2, 1, 0)]
            out = np.array(ket)
            for a in range(2 ** len(sel)):
                val = {c: (a >> i) & 1 for i, c in enumerate(sel)}
                u = _phase_unitary(o, coefs, srcs, val)
                # the slice of the state whose selector bits take these values
                idx = [slice(None)] * n
                for c, x in val.items():
                    idx[n - 1 - index_bit(c)] = x
                kept = [ax for ax in range(n) if isinstance(idx[ax], slice)]
                sub = O.tensormul(u.reshape([2] * 8), ket[tuple(idx)],
                                  [kept.index(ax) for ax in axes])
                out[tuple(idx)] = sub
            ket = out
        else:
            k = o.k
            axes = [n - 1 - pos[o.t[j]] for j in range(k)]
            if o.kind == fusion.OP_DENSE:
                m = coefs[o.coef:o.coef + 4 ** k].reshape(2 ** k, 2 ** k)
            elif o.kind == fusion.OP_MFMA:
                m = unpack_mfma(coefs[o.coef:o.coef + 2 * 4 ** k], k)
            else:
                m = np.zeros((2 ** k, 2 ** k), dtype=complex)
                for r in range(2 ** k):
                    m[r, srcs[o.src + r]] = coefs[o.coef + r]
            ket = O.tensormul(m.reshape([2] * (2 * k)), ket, axes)
    return ket


@pytest.mark.parametrize("T,seed", [(7, 0), (8, 1), (10, 2)])
def test_register_phase_encoding_replays(T, seed):
    n = 11
    desc = random_mixed_circuit(n, 10, seed=seed)
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    bits_list = [tuple(n - 1 - q for q in g.qubits) for g in gates]
    ket = O.zero_state(n)
    nphase = 0
    for pos, members in fusion.plan_passes(bits_list, n, T, 3):
        pidx = {b: j for j, b in enumerate(pos)}
        coefs, srcs, ops = [], [], []
        for item in fusion.plan_phases([fusion.Op(bits_list[i], gates[i].kind) for i in members]):
            if item[0] == 'phase':
                assert len(item[1]) <= fusion.REG_BITS
                ops.append(fusion._encode_phase(item[1], item[2], pidx, T, coefs, srcs))
            else:
                ops.append(fusion._encode_op(item[1], pidx, coefs, srcs))
        nphase += len(ops)
        fusion.prepare_ops(ops)
        srcs.extend([0, 0, 0, 0])           # the record read ahead (as plans do)
        arr = (_native.QfxTileOp * len(ops))(*ops)
        assert _native.load().qfx_tile_check_ops(T, 256, 0, arr, len(ops), len(coefs),
                                                 _native.int_array(srcs), len(srcs)) == 0
        ket = _emulate_ops(ket, n, pos, ops, coefs, srcs)
    assert nphase < len(gates)
    assert np.max(np.abs(ket - ref)) < 1e-12


def test_exchange_patterns():
    from quantumflow_amd import stdgates as sg
    assert fusion.exchange_pattern(sg.X().kind) == ('cx', 0, [])
    assert fusion.exchange_pattern(sg.CNOT().kind) == ('cx', 1, [0])
    assert fusion.exchange_pattern(sg.CCNOT().kind) == ('cx', 2, [0, 1])
    assert fusion.exchange_pattern(sg.SWAP().kind) == ('swap', 0, 1, [])
    assert fusion.exchange_pattern(sg.CSWAP().kind) == ('swap', 1, 2, [0])
    assert fusion.exchange_pattern(sg.ISWAP().kind) == ('swap', 0, 1, [])
    assert fusion.exchange_pattern(sg.Y().kind) == ('cx', 0, [])
    assert fusion.exchange_pattern(sg.H().kind) is None


def test_uncontrolled_exchanges_cost_no_sub_ops():
    """SWAP, X and ISWAP's permutation become relabelings: a phase of only
    SWAP/X/ISWAP gates holds at most the merged ISWAP phases."""
    n, T = 8, 8
    desc = [('SWAP', (), (0, 1)), ('X', (), (2,)), ('ISWAP', (), (1, 2)), ('SWAP', (), (2, 3)),
            ('X', (), (0,)), ('ISWAP', (), (0, 3))]
    gates = build_gates(desc)
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                           fuse_width=1, mode='reg')
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    ket = O.zero_state(n)
    rng = np.random.default_rng(3)
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    psi = (psi / np.linalg.norm(psi)).reshape([2] * n)
    ref = O.run_circuit(desc, qubits=tuple(range(n)), ket=psi)
    ket = psi
    nsub = 0
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        ops = plan.ops_host[start:start + nops]
        for o in ops:
            nsub += o.coef
            for q in range(o.coef):
                assert plan.srcs[o.src + 4 * q] == fusion.SUB_DIAG
        ket = _emulate_ops(ket, n, pos, ops, plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert nsub <= 2
    assert np.array_equal(np.abs(ket), np.abs(ref))
    assert np.max(np.abs(ket - ref)) < 1e-15


@pytest.mark.parametrize("seed", [5, 6])
def test_register_phases_permutation_heavy(seed):
    """Circuits of X/CNOT/SWAP/ISWAP/CCNOT/CSWAP/Y and phases replay exactly."""
    rng = np.random.default_rng(seed)
    n, T = 9, 7
    names = [('X', 1), ('Y', 1), ('CNOT', 2), ('SWAP', 2), ('ISWAP', 2), ('CCNOT', 3),
             ('CSWAP', 3), ('CZ', 2), ('S', 1), ('H', 1)]
    desc = []
    for _ in range(80):
        nm, k = names[int(rng.integers(len(names)))]
        desc.append((nm, (), tuple(int(q) for q in rng.permutation(n)[:k])))
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    bits_list = [tuple(n - 1 - q for q in g.qubits) for g in gates]
    ket = O.zero_state(n)
    for pos, members in fusion.plan_passes(bits_list, n, T, 3):
        pidx = {b: j for j, b in enumerate(pos)}
        coefs, srcs, ops = [], [], []
        for item in fusion.plan_phases([fusion.Op(bits_list[i], gates[i].kind) for i in members]):
            if item[0] == 'phase':
                ops.append(fusion._encode_phase(item[1], item[2], pidx, T, coefs, srcs))
            else:
                ops.append(fusion._encode_op(item[1], pidx, coefs, srcs))
        fusion.prepare_ops(ops)
        srcs.extend([0, 0, 0, 0])           # the record read ahead (as plans do)
        arr = (_native.QfxTileOp * len(ops))(*ops)
        assert _native.load().qfx_tile_check_ops(T, 256, 0, arr, len(ops), len(coefs),
                                                 _native.int_array(srcs), len(srcs)) == 0
        ket = _emulate_ops(ket, n, pos, ops, coefs, srcs)
    assert np.max(np.abs(ket - ref)) < 1e-12


def mfma_16x16x4(a_lanes, b_lanes, acc):
    """v_mfma_f64_16x16x4f64 on 64 lanes by its operand layout: lane l holds
    A[l % 16][l // 16], B[l // 16][l % 16] and D[4 j + l // 16][l % 16] in
    register j."""
    A = np.zeros((16, 4))
    B = np.zeros((4, 16))
    for l in range(64):
        A[l % 16, l // 16] = a_lanes[l]
        B[l // 16, l % 16] = b_lanes[l]
    D = A @ B
    out = acc.copy()
    for l in range(64):
        for j in range(4):
            out[l, j] += D[4 * j + l // 16, l % 16]
    return out


@pytest.mark.parametrize("k", [3, 4, 5])
def test_mfma_block_lane_program(k):
    """The matrix-core op's lane program (which amplitudes a lane reads, the
    packed operands it feeds, which result registers it writes back), run
    on an emulated 16x16x4 instruction, computes M x for 16 groups."""
    rng = np.random.default_rng(k)
    d = 1 << k
    m = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    x = rng.normal(size=(16, d)) + 1j * rng.normal(size=(16, d))    # group, amplitude
    p = fusion.pack_mfma(m)
    na, nk, nr = d // 4, 2 * d // 4, 2 * d // 16
    lanes = np.arange(64)
    q, c = lanes // 16, lanes % 16
    xs = [x[c, 4 * s + q] for s in range(na)]                  # x[s] per lane
    b = [xs[ks].real if ks < na else xs[ks - na].imag for ks in range(nk)]
    y = np.zeros((16, d), dtype=complex)
    if k == 3:
        acc = np.zeros((64, 4))
        for ks in range(nk):
            acc = mfma_16x16x4(p[ks * 64:(ks + 1) * 64], b[ks], acc)
        for s in range(2):
            y[c, 4 * s + q] = acc[:, s] + 1j * acc[:, s + 2]
    else:
        npair = nr // 2
        for rp in range(npair):
            re = np.zeros((64, 4))
            im = np.zeros((64, 4))
            for ks in range(nk):
                re = mfma_16x16x4(p[(rp * nk + ks) * 64:][:64], b[ks], re)
                im = mfma_16x16x4(p[((rp + npair) * nk + ks) * 64:][:64], b[ks], im)
            for j in range(4):
                y[c, 4 * (4 * rp + j) + q] = re[:, j] + 1j * im[:, j]
    assert np.max(np.abs(y - x @ m.T)) < 1e-12
    assert np.array_equal(unpack_mfma(p.view(np.complex128), k), m)


@pytest.mark.parametrize("width,T", [(3, 8), (4, 9), (5, 10)])
def test_mfma_plan_replays(width, T):
    """mode='mfma' plans: every block holds 3..width qubits, the encoded
    blocks validate, and replaying them gives the circuit's state."""
    n = 11
    desc = random_mixed_circuit(n, 10, seed=width)
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                           fuse_width=width, mode='mfma')
    assert plan.mode == 'mfma' and plan.n_blocks < len(gates)
    # matrix-core passes read their coefficients from global memory: never cut
    nbits = [tuple(n - 1 - q for q in g.qubits) for g in gates]
    assert plan.n_passes == len(fusion.plan_passes(nbits, n, T, 3))
    ket = O.zero_state(n)
    for kind, payload in plan.passes:
        assert kind == 'tile'
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        assert flags & fusion.TILE_MFMA
        ops = plan.ops_host[start:start + nops]
        assert all(o.kind == fusion.OP_MFMA and 3 <= o.k <= width for o in ops)
        ket = _emulate_ops(ket, n, pos, ops, plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert np.max(np.abs(ket - ref)) < 1e-12


def test_mfma_validator_rules():
    """Matrix-core blocks need QFX_TILE_MFMA, 3 <= k <= 5, T >= k + 4 and
    their 2 * 4^k coefficient slots; 4-5 qubit dense ops need QFX_TILE_WIDE."""
    lib = _native.load()
    coefs = []
    pidx = {b: b for b in range(9)}
    o = fusion._encode_mfma([0, 1, 2, 3], np.eye(16), pidx, coefs)
    arr = (_native.QfxTileOp * 1)(o)
    srcs = _native.int_array([0])
    M = fusion.TILE_MFMA
    assert lib.qfx_tile_check_ops(9, 256, M, arr, 1, len(coefs), srcs, 1) == 0
    assert lib.qfx_tile_check_ops(9, 256, 0, arr, 1, len(coefs), srcs, 1) == _native.QFX_EINVAL
    assert lib.qfx_tile_check_ops(7, 256, M, arr, 1, len(coefs), srcs, 1) == _native.QFX_EINVAL
    assert lib.qfx_tile_check_ops(9, 256, M, arr, 1, len(coefs) - 1, srcs, 1) == _native.QFX_EINVAL
    assert lib.qfx_tile_check_ops(9, 1024, M, arr, 1, len(coefs), srcs, 1) == _native.QFX_EINVAL
    dense = fusion._encode_op(fusion.Op([0, 1, 2, 3], None, np.eye(16)), pidx, coefs, [])
    arr = (_native.QfxTileOp * 1)(dense)
    assert lib.qfx_tile_check_ops(9, 256, 0, arr, 1, len(coefs), srcs, 1) == _native.QFX_EINVAL
    assert lib.qfx_tile_check_ops(9, 256, fusion.TILE_WIDE, arr, 1, len(coefs), srcs, 1) == 0


def test_d1_variants():
    from quantumflow_amd import stdgates as sg
    assert fusion.d1_variant(sg.RY(0.3).kind.matrix) == fusion.SUB_D1R
    assert fusion.d1_variant(sg.H().kind.matrix) == fusion.SUB_D1R
    assert fusion.d1_variant(sg.RX(0.3).kind.matrix) == fusion.SUB_D1X
    assert fusion.d1_variant(sg.RX(0.3).kind.matrix @ sg.RY(0.2).kind.matrix) == fusion.SUB_D1
    assert fusion.d1_variant(sg.T().kind.matrix) == fusion.SUB_D1


@pytest.mark.parametrize("mode,width", [('reg', 1), ('lds', 3)])
def test_host_plan_pass_tables_replay(mode, width):
    """Each pass indexes its own coefficient table (what the kernel stages in
    LDS); passes over the staging limit are cut; replaying every pass with
    its table gives the circuit's state."""
    n = 12
    desc = random_mixed_circuit(n, 14, seed=11)
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=9, low_bits=3, block=256,
                           fuse_width=width, mode=mode)
    ket = O.zero_state(n)
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        assert ncoef <= fusion.MAX_STAGED or flags & (fusion.TILE_WIDE | fusion.TILE_MFMA)
        ket = _emulate_ops(ket, n, pos, plan.ops_host[start:start + nops],
                           plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert np.max(np.abs(ket - ref)) < 1e-12


def test_pass_split_at_staging_limit(monkeypatch):
    """A pass whose coefficients exceed the staging limit is cut into
    consecutive passes on the same tile bits."""
    monkeypatch.setattr(fusion, 'MAX_STAGED', 40)
    n = 10
    desc = random_mixed_circuit(n, 8, seed=2)
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=8, low_bits=3, block=256,
                           fuse_width=1, mode='reg')
    unsplit = fusion.plan_passes([tuple(n - 1 - q for q in g.qubits) for g in gates], n, 8, 3)
    assert plan.n_passes > len(unsplit)
    ket = O.zero_state(n)
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        assert ncoef <= 40 or nops == 1
        ket = _emulate_ops(ket, n, pos, plan.ops_host[start:start + nops],
                           plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert np.max(np.abs(ket - ref)) < 1e-12


@pytest.mark.parametrize("virtual", [False, True])
def test_pass_level_swap_relabel(virtual):
    """SWAP/ISWAP gates in a pass become a relabelling of tile bits plus the
    ISWAP phases; the pass stores through a permutation; replaying with the
    store permutation gives the circuit's state exactly.  With virtual swaps
    the SWAP-type gates are renamed away before planning and restored by
    swaps at the end; the replay is exact as well."""
    n, T = 10, 8
    rng = np.random.default_rng(12)
    names = [('SWAP', 2), ('ISWAP', 2), ('RX', 1), ('CNOT', 2), ('CZ', 2), ('H', 1)]
    desc = []
    for _ in range(40):
        nm, k = names[int(rng.integers(len(names)))]
        params = (float(rng.uniform(0, 6)),) if nm == 'RX' else ()
        desc.append((nm, params, tuple(int(q) for q in rng.permutation(n)[:k])))
    gates = build_gates(desc)
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    psi = (psi / np.linalg.norm(psi)).reshape([2] * n)
    ref = O.run_circuit(desc, qubits=tuple(range(n)), ket=psi)
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                           fuse_width=1, mode='reg', virtual_swaps=virtual)
    if virtual:
        assert plan.n_virtual == sum(d[0] in ('SWAP', 'ISWAP') for d in desc) > 0
    else:
        assert any(p[1][7] is not None for p in plan.passes)
    ket = psi
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        if spos is not None:
            # tile bits permute among themselves, the others and the three
            # lowest index bits (the coalesced runs) never move
            assert sorted(spos) == list(range(n)) and all(spos[b] == b for b in range(3))
            assert all(spos[b] == b for b in range(n) if b not in pos)
            assert sorted(spos[b] for b in pos) == sorted(pos)
        ket = _emulate_ops(ket, n, pos, plan.ops_host[start:start + nops],
                           plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert np.max(np.abs(ket - ref)) < 1e-12


def _sw(e):
    # tile element e lives at LDS slot sw(e): bits 3.., in groups of three,
    # folded into bits 0..2
    return e ^ (((e >> 3) ^ (e >> 6) ^ (e >> 9) ^ (e >> 12)) & 7)


def test_prepared_lds_words_and_stale_words_rejected():
    """qfx_tile_prepare_ops derives, per register phase, the LDS slot offset
    of each load/store map column and constant: sw() of the tile-bit mask
    the column selects.  The validator rejects a phase whose words do not
    match its bits and maps."""
    desc = random_mixed_circuit(10, 8, seed=4)
    gates = build_gates(desc)
    plan = fusion.HostPlan(gates, tuple(range(10)), tile_bits=8, low_bits=3, block=256,
                           fuse_width=1, mode='reg', persistent=False)
    phases = [o for o in plan.ops_host if o.kind == fusion.OP_RPHASE]
    assert phases
    for o in phases:
        t = [o.t[j] for j in range(fusion.REG_BITS)]

        def slot(x):
            return _sw(sum(1 << t[j] for j in range(4) if (x >> j) & 1))
        want = [slot((o.phase >> (4 * j)) & 15) for j in range(4)] + [slot(o.phase >> 16)]
        want += [slot((o.active >> (4 * j)) & 15) for j in range(4)] + [slot(o.active >> 16)]
        assert list(o.lds) == want
    lib = _native.load()
    o = phases[0]
    arr = (_native.QfxTileOp * 1)(o)
    srcs = _native.int_array(plan.srcs)
    ncoef = len(plan.coefs)
    assert lib.qfx_tile_check_ops(8, 256, 0, arr, 1, ncoef, srcs, len(plan.srcs)) == 0
    arr[0].lds[9] ^= 1
    assert lib.qfx_tile_check_ops(8, 256, 0, arr, 1, ncoef, srcs,
                                  len(plan.srcs)) == _native.QFX_EINVAL
    assert lib.qfx_tile_prepare_ops(arr, 1) == 0
    assert lib.qfx_tile_check_ops(8, 256, 0, arr, 1, ncoef, srcs, len(plan.srcs)) == 0


@pytest.mark.parametrize("n,T,seed", [(12, 8, 0), (12, 9, 1), (13, 8, 2)])
def test_selector_diagonals_replay(n, T, seed):
    """Diagonal gates join the first pass that reaches them whatever its tile
    bits; their phases on index bits outside the tile become selector
    diagonals (SUB_DIAGV, up to 3 such bits per table).  Passes drop, and
    replaying the plan gives the circuit's state."""
    rng = np.random.default_rng(seed)
    desc = []
    for _ in range(8):
        for q in range(n):
            desc.append((('RZ', 'PHASE', 'RX')[rng.integers(3)], (float(rng.uniform(0, 6)),), (q,)))
        order = rng.permutation(n)
        for j in range(0, n - 1, 2):
            a, b = int(order[j]), int(order[j + 1])
            nm = ('CZ', 'CNOT', 'CPHASE')[rng.integers(3)]
            desc.append((nm, (float(rng.uniform(0, 6)),) if nm == 'CPHASE' else (), (a, b)))
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                           fuse_width=1, mode='reg')
    nsel = ncompact = 0
    ket = O.zero_state(n)
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        ops = plan.ops_host[start:start + nops]
        for o in ops:
            if o.kind == fusion.OP_RPHASE:
                for q in range(o.coef):
                    if plan.srcs[o.src + 4 * q] == fusion.SUB_DIAGV:
                        nsel += 1
                        ncompact += plan.srcs[o.src + 4 * q + 3] != 0
        ket = _emulate_ops(ket, n, pos, ops, plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert nsel > 0 and ncompact > 0
    assert np.max(np.abs(ket - ref)) < 1e-12
    # without selector diagonals the same circuit needs more passes
    bits_list = [tuple(n - 1 - q for q in g.qubits) for g in gates]
    plain = fusion.plan_passes(bits_list, n, T, 3)
    assert plan.n_passes < len(plain)


def test_selector_diagonal_validator_rules():
    """SUB_DIAGV records: 1..3 selector bits, index bits < 64, no stray code
    bits, and 16 * 2^m coefficients in range."""
    lib = _native.load()
    o = _native.QfxTileOp()
    o.kind = fusion.OP_RPHASE
    o.k = 4
    for j in range(4):
        o.t[j] = o.sorted[j] = j
    o.coef, o.src = 1, 0
    o.active = o.phase = sum(1 << (5 * j) for j in range(4))   # identity maps
    fusion.prepare_ops([o])
    arr = (_native.QfxTileOp * 1)(o)

    def check(code, ncoef):
        srcs = _native.int_array([fusion.SUB_DIAGV, code, 0, 0])
        return lib.qfx_tile_check_ops(8, 256, 0, arr, 1, ncoef, srcs, 4)
    assert check(2 | (20 << 2) | (33 << 9), 64) == 0
    assert check(2 | (20 << 2) | (33 << 9), 63) == _native.QFX_EINVAL
    assert check(0, 16) == _native.QFX_EINVAL
    assert check(1 | (5 << 2) | (1 << 9), 32) == _native.QFX_EINVAL      # stray bits
    assert check(3 | (10 << 2) | (11 << 9) | (12 << 16), 128) == 0
    # compact tables (fourth word 16 | d): 2^popcount(d) phases per table
    def check_c(code, w, ncoef):
        srcs = _native.int_array([fusion.SUB_DIAGV, code, 0, w])
        return lib.qfx_tile_check_ops(8, 256, 0, arr, 1, ncoef, srcs, 4)
    assert check_c(2 | (20 << 2) | (33 << 9), 16 | 0b0101, 16) == 0
    assert check_c(2 | (20 << 2) | (33 << 9), 16 | 0b0101, 15) == _native.QFX_EINVAL
    assert check_c(1 | (20 << 2), 16, 2) == 0                        # no register bit
    assert check_c(1 | (20 << 2), 32 | 1, 64) == _native.QFX_EINVAL  # stray flag bits
    # flip-mask bits the tables read (mask bits 8..11): no more coefficients
    assert check_c(1 | (20 << 2), 16 | 0b0101 | 0b0100 << 8, 8) == 0
    assert check_c(1 | (20 << 2), 16 | 0b0101 | 0b0100 << 8, 7) == _native.QFX_EINVAL
    assert check_c(0, 0b0011 << 8, 16) == 0                          # flip bits only
    assert check_c(0, 0b0011 << 8, 15) == _native.QFX_EINVAL
    assert check_c(0, 0, 16) == _native.QFX_EINVAL                   # no selector
    assert check_c(1 | (20 << 2), 0b0101, 64) == _native.QFX_EINVAL  # dep without 16
    assert check_c(1 | (20 << 2), 1 << 12, 64) == _native.QFX_EINVAL
    # lane-group selectors: tile bits outside the registers (here t = 0..3)
    assert check(1 | ((64 + 5) << 2), 32) == 0
    assert check(1 | ((64 + 2) << 2), 32) == _native.QFX_EINVAL         # a register bit
    assert check(1 | ((64 + 8) << 2), 32) == _native.QFX_EINVAL         # beyond T = 8


@pytest.mark.parametrize("n,T,seed", [(12, 8, 0), (13, 9, 1), (12, 7, 2), (14, 8, 3)])
def test_selector_controls_replay(n, T, seed):
    """Controlled X gates need only their target in the tile and in the
    phase's registers: their other controls select per lane group (tile
    bits outside the registers) or per tile (index bits outside the tile)
    whether they act (SUB_CX selector controls).  Replaying the plan gives
    the circuit's state, and the circuit takes fewer passes than with every
    control in the tile."""
    rng = np.random.default_rng(seed)
    desc = []
    for _ in range(6):
        for q in rng.permutation(n)[:n // 2]:
            nm = ('RX', 'RY', 'H', 'T')[rng.integers(4)]
            desc.append((nm, (float(rng.uniform(0, 6)),) if nm in ('RX', 'RY') else (), (int(q),)))
        for _ in range(n // 2):
            nm = ('CNOT', 'CCNOT', 'CZ', 'CNOT')[rng.integers(4)]
            k = 3 if nm == 'CCNOT' else 2
            desc.append((nm, (), tuple(int(x) for x in rng.choice(n, k, replace=False))))
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                           fuse_width=1, mode='reg')
    lane = glob = 0
    ket = O.zero_state(n)
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        ops = plan.ops_host[start:start + nops]
        for o in ops:
            if o.kind != fusion.OP_RPHASE:
                continue
            for q in range(o.coef):
                typ, w = plan.srcs[o.src + 4 * q], plan.srcs[o.src + 4 * q + 3]
                f = typ >> 8                # a lane flip
                codes = [f >> 3 & 127] if f else []
                codes += [f >> 12 & 127] if f & 2048 else []
                if typ & fusion.REC_TYPE == fusion.SUB_CX and w:
                    codes += [(w >> (2 + 7 * j)) & 127 for j in range(w & 3)]
                if codes:
                    lane += any(c >= 64 for c in codes)
                    glob += any(c < 64 for c in codes)
        ket = _emulate_ops(ket, n, pos, ops, plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert lane > 0 and glob > 0
    assert np.max(np.abs(ket - ref)) < 1e-12
    plain = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                            fuse_width=1, mode='reg', free_controls=False)
    assert plan.n_passes < plain.n_passes


def _lane_flip_replay(seed, flips):
    """Replay of a plan of a 14-qubit random circuit (max deviation from the
    oracle's state) and its counts of flips and flagged records."""
    n, T = 14, 9
    desc = random_mixed_circuit(n, 14, seed=seed)
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                           fuse_width=1, mode='reg', lane_flips=flips)
    ket = O.zero_state(n)
    count = {'flip': 0, 'd1': 0, 'cx': 0, 'past': 0}
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        ops = plan.ops_host[start:start + nops]
        for o in ops:
            for q in range(o.coef if o.kind == fusion.OP_RPHASE else 0):
                typ, code = plan.srcs[o.src + 4 * q: o.src + 4 * q + 2]
                count['flip'] += bool(typ >> 8)
                # a flip riding past the record before it (that record
                # carries one too)
                count['past'] += bool(typ >> 8) and q > 0 and bool(plan.srcs[o.src + 4 * q - 4] >> 8)
                typ &= fusion.REC_TYPE
                count['d1'] += (typ in (fusion.SUB_D1, fusion.SUB_D1R, fusion.SUB_D1X)
                                and bool(code & fusion.D1_FLIP))
                count['cx'] += typ == fusion.SUB_CX and bool(code & fusion.CX_FLIP)
        ket = _emulate_ops(ket, n, pos, ops, plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    return np.max(np.abs(ket - ref)), count


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_lane_flips_replay(seed):
    """A controlled X whose controls are all selectors becomes a lane flip
    riding on the next record: no amplitudes move, later diagonals take it
    into their selector tables, 2x2 matrices and CX gates reading its bit
    carry their flag, and the phase's store puts every register where it
    belongs.  Replayed, plans with and without flips give the circuit's
    state."""
    err, count = _lane_flip_replay(seed, True)
    assert err < 1e-12 and count['flip'] > 0
    err, count = _lane_flip_replay(seed, False)
    assert err < 1e-12 and not any(count.values())


def test_lane_flip_flags_all_seen():
    """Over those circuits the plans hold flips on consecutive records and
    flags of every kind, so the replays above cover each flagged reading."""
    kinds = set()
    for seed in range(4):
        kinds |= {k for k, v in _lane_flip_replay(seed, True)[1].items() if v}
    assert kinds == {'flip', 'd1', 'cx', 'past'}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_store_map_cnots_replay(seed):
    """A CNOT with control and target in registers that commutes with every
    later sub-op of its phase is absorbed by the phase's affine store map,
    one that commutes with every earlier sub-op by its load map (diagonals
    it passes re-indexed): fewer exchange records, same state."""
    n, T = 12, 9
    desc = random_mixed_circuit(n, 20, seed=40 + seed)
    gates = build_gates(desc)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    ncx = {}
    for store_cx in (True, False):
        plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                               fuse_width=1, mode='reg', store_cx=store_cx)
        ket = O.zero_state(n)
        for kind, payload in plan.passes:
            pos, start, nops, block, flags, c_lo, ncoef, spos = payload
            ops = plan.ops_host[start:start + nops]
            ket = _emulate_ops(ket, n, pos, ops, plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
            ket = apply_store(ket, n, pos, spos)
        assert np.max(np.abs(ket - ref)) < 1e-12
        ncx[store_cx] = sum(plan.srcs[o.src + 4 * q] & fusion.REC_TYPE == fusion.SUB_CX
                            for o in plan.ops_host if o.kind == fusion.OP_RPHASE
                            for q in range(o.coef))
    assert ncx[True] < ncx[False]


def test_same_bit_matrices_merge_across_other_bits():
    """Two 2x2 matrices on one register bit with only records that leave
    that bit alone between them become one record (their product)."""
    import quantumflow_amd as qf
    n = 10
    gates = [qf.RX(0.3, 9), qf.CZ(8, 7), qf.RY(0.2, 6), qf.RY(0.7, 9), qf.H(8)]
    desc = [('RX', (0.3,), (9,)), ('CZ', (), (8, 7)), ('RY', (0.2,), (6,)),
            ('RY', (0.7,), (9,)), ('H', (), (8,))]
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=8, low_bits=3, block=256,
                           fuse_width=1, mode='reg')
    recs = [plan.srcs[o.src + 4 * q] & fusion.REC_TYPE for o in plan.ops_host
            if o.kind == fusion.OP_RPHASE for q in range(o.coef)]
    two_by_two = [t for t in recs if t in (fusion.SUB_D1, fusion.SUB_D1R, fusion.SUB_D1X)]
    assert len(two_by_two) == 3             # RX and RY on qubit 9 merged
    ket = O.zero_state(n)
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        ket = _emulate_ops(ket, n, pos, plan.ops_host[start:start + nops],
                           plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert np.max(np.abs(ket - ref)) < 1e-12


def test_diagonals_merge_across_other_records():
    """A diagonal moves forward past records that leave its bits alone (here
    a CNOT controlled outside the tile, whose selector control made the
    planner emit the first diagonal early) to the next diagonal, and the
    two become one record; without merging they stay apart."""
    import quantumflow_amd as qf
    n = 10
    desc = [('CZ', (), (9, 8)), ('CNOT', (), (0, 7)), ('PHASE', (0.4,), (9,)),
            ('RY', (0.5,), (6,))]
    gates = [qf.CZ(9, 8), qf.CNOT(0, 7), qf.PHASE(0.4, 9), qf.RY(0.5, 6)]
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    for merge in (False, True):
        plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=8, low_bits=3, block=256,
                               fuse_width=1, mode='reg', store_cx=False, merge_records=merge)
        recs = [plan.srcs[o.src + 4 * q] & fusion.REC_TYPE for o in plan.ops_host
                if o.kind == fusion.OP_RPHASE for q in range(o.coef)]
        # merged, the two diagonals (on lane-group bits only) become one, and
        # that rides in the RY record as one matrix per selector value
        assert sum(t in (fusion.SUB_DIAG, fusion.SUB_DIAGV) for t in recs) == (0 if merge else 2)
        assert len(recs) == (1 if merge else 3)
        ket = O.zero_state(n)
        for kind, payload in plan.passes:
            pos, start, nops, block, flags, c_lo, ncoef, spos = payload
            ket = _emulate_ops(ket, n, pos, plan.ops_host[start:start + nops],
                               plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
            ket = apply_store(ket, n, pos, spos)
        assert np.max(np.abs(ket - ref)) < 1e-12


def test_lane_flip_validator_rules():
    """A lane-flip word riding on a record (type bits 8..): bit 0 set,
    selectors as for SUB_DIAGV, nothing above bit 19; flagged 2x2 records
    need 8 coefficients."""
    T = 9
    pidx = {b: b for b in range(T)}
    coefs, srcs = [], []
    o = fusion._encode_phase([0, 1, 2, 3], [], pidx, T, coefs, srcs)
    lib = _native.load()

    def check(rec, ncoef=16):
        o.src, o.coef = 0, 1
        fusion.prepare_ops([o])
        arr = (_native.QfxTileOp * 1)(o)
        return lib.qfx_tile_check_ops(T, 256, 0, arr, 1, ncoef, _native.int_array(rec), 4)
    f = 1 | (2 << 1) | (70 << 3) | (1 << 10)        # bit 2 where tile bit 6 is 1
    g = f | (1 << 11) | (40 << 12)                   # and index bit 40 is 0
    d1 = [fusion.SUB_D1R, 3, 0, 0]
    assert check(d1) == 0
    assert check([d1[0] | f << 8, 3, 0, 0]) == 0
    assert check([d1[0] | g << 8, 3, 4, 0]) == 0
    assert check([d1[0] | (f ^ 1) << 8, 3, 0, 0]) == _native.QFX_EINVAL      # bit 0 clear
    assert check([d1[0] | (f | 1 << 20) << 8, 3, 0, 0]) == _native.QFX_EINVAL
    assert check([d1[0] | (f | 1 << 12) << 8, 3, 0, 0]) == _native.QFX_EINVAL  # stray selector
    bad = 1 | (66 << 3)                              # a register bit as selector
    assert check([d1[0] | bad << 8, 3, 0, 0]) == _native.QFX_EINVAL
    assert check([fusion.SUB_D1R, 3 | fusion.D1_FLIP, 8, 0]) == 0
    assert check([fusion.SUB_D1R, 3 | fusion.D1_FLIP, 9, 0]) == _native.QFX_EINVAL
    assert check([fusion.SUB_D1R, 8, 0, 0]) == _native.QFX_EINVAL
    assert check([fusion.SUB_CX, 1 | (1 << 2) | (1 << 6) | fusion.CX_FLIP, 0, 0]) == 0
    assert check([fusion.SUB_CX, 1 << 11, 0, 0]) == _native.QFX_EINVAL
    assert check([fusion.SUB_DIAG, 1, 0, 0]) == _native.QFX_EINVAL


def test_selector_control_validator_rules():
    """SUB_CX selector controls (the record's fourth word): none, or 1..3
    selectors coded as for SUB_DIAGV with their required values in bits
    23 + j, nothing else."""
    lib = _native.load()
    o = _native.QfxTileOp()
    o.kind = fusion.OP_RPHASE
    o.k = 4
    for j in range(4):
        o.t[j] = o.sorted[j] = j
    o.coef, o.src = 1, 0
    o.active = o.phase = sum(1 << (5 * j) for j in range(4))   # identity maps
    fusion.prepare_ops([o])
    arr = (_native.QfxTileOp * 1)(o)

    def check(w, code=1 | (1 << 2) | (1 << 6)):     # X on bit 1 controlled by bit 0
        srcs = _native.int_array([fusion.SUB_CX, code, 0, w])
        return lib.qfx_tile_check_ops(8, 256, 0, arr, 1, 0, srcs, 4)
    assert check(0) == 0
    assert check(1 | (20 << 2) | (1 << 23)) == 0                      # index bit 20 on |1>
    assert check(1 | (20 << 2)) == 0                                  # ... on |0>
    assert check(2 | (20 << 2) | ((64 + 5) << 9) | (3 << 23)) == 0    # and lane bit 5
    assert check(3 | (10 << 2) | (11 << 9) | (12 << 16) | (7 << 23)) == 0
    assert check(1 << 23) == _native.QFX_EINVAL                       # values, no selector
    assert check(1 | (20 << 2) | (1 << 24)) == _native.QFX_EINVAL     # value beyond m
    assert check(1 | (20 << 2) | (5 << 9)) == _native.QFX_EINVAL      # stray selector bits
    assert check(1 | ((64 + 2) << 2)) == _native.QFX_EINVAL           # a register bit
    assert check(1 | ((64 + 8) << 2)) == _native.QFX_EINVAL           # beyond T = 8
    assert check(0, code=1 | (2 << 2)) == _native.QFX_EINVAL          # control on the target


def test_gate_types_commutation_brute_force():
    """Every pair of standard gates on 3 qubits whose shared qubits all carry
    the same 'Z' or 'X' type (fusion.gate_types) commutes, checked on the
    full 8x8 operators; pairs that gate_types does not clear include
    non-commuting ones (the rule is not vacuous)."""
    from quantumflow_amd import stdgates as S
    rng = np.random.default_rng(5)
    makers = [lambda q: S.RX(float(rng.uniform(0, 6)), q[0]),
              lambda q: S.RY(float(rng.uniform(0, 6)), q[0]),
              lambda q: S.RZ(float(rng.uniform(0, 6)), q[0]),
              lambda q: S.PHASE(float(rng.uniform(0, 6)), q[0]),
              lambda q: S.X(q[0]), lambda q: S.H(q[0]),
              lambda q: S.CNOT(q[0], q[1]), lambda q: S.CZ(q[0], q[1]),
              lambda q: S.ISWAP(q[0], q[1]), lambda q: S.SWAP(q[0], q[1]),
              lambda q: S.CPHASE(float(rng.uniform(0, 6)), q[0], q[1]),
              lambda q: S.CCNOT(q[0], q[1], q[2])]
    arity = [1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 3]

    def full(g):
        qs = list(g.qubits)
        return O.full_operator(g.matrix.reshape([2] * (2 * len(qs))), 3, qs)

    cleared = blocked_noncomm = 0
    for _ in range(600):
        a, b = rng.integers(len(makers), size=2)
        qa = list(rng.permutation(3)[:arity[a]])
        qb = list(rng.permutation(3)[:arity[b]])
        ga, gb = makers[a]([int(x) for x in qa]), makers[b]([int(x) for x in qb])
        ta = fusion.gate_types(ga.kind, [2 - q for q in ga.qubits])
        tb = fusion.gate_types(gb.kind, [2 - q for q in gb.qubits])
        shared = set(ta) & set(tb)
        ma, mb = full(ga), full(gb)
        comm = np.max(np.abs(ma @ mb - mb @ ma))
        if all(ta[x] == tb[x] != 'G' for x in shared):
            cleared += 1
            assert comm < 1e-12, (ga, gb, ta, tb)
        elif comm > 1e-6:
            blocked_noncomm += 1
    assert cleared > 50 and blocked_noncomm > 50


def test_pass_of_only_absorbed_swaps_keeps_its_store_permutation():
    """A pass whose gates are all absorbed exchanges has no ops left: it must
    still run, as a pass that only stores through the permutation."""
    n, T = 10, 8
    desc = [('SWAP', (), (0, 1)), ('SWAP', (), (1, 2))]
    gates = build_gates(desc)
    rng = np.random.default_rng(9)
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    psi = (psi / np.linalg.norm(psi)).reshape([2] * n)
    ref = O.run_circuit(desc, qubits=tuple(range(n)), ket=psi)
    plan = fusion.HostPlan(gates, tuple(range(n)), tile_bits=T, low_bits=3, block=256,
                           fuse_width=1, mode='reg')
    assert any(p[1][2] == 0 and p[1][7] is not None for p in plan.passes)
    ket = psi
    for kind, payload in plan.passes:
        pos, start, nops, block, flags, c_lo, ncoef, spos = payload
        ket = _emulate_ops(ket, n, pos, plan.ops_host[start:start + nops],
                           plan.coefs[c_lo:c_lo + ncoef], plan.srcs)
        ket = apply_store(ket, n, pos, spos)
    assert np.array_equal(ket, ref)


def test_plan_flops_closed_form():
    """HostPlan.flops: 2^n times the per-amplitude FP64 operations of each
    sub-op: a real or RX-type 2x2 matrix 6, a general 2x2 matrix 14 (two
    complex multiplies and an add per output), a phase 6 (one complex
    multiply), an exchange 0."""
    import quantumflow_amd as qf
    n = 8
    cases = [([qf.H(0)], 6), ([qf.RX(0.3, 1)], 6), ([qf.RZ(0.3, 2)], 6), ([qf.CNOT(0, 5)], 0),
             ([qf.TX(0.3, 1)], 14), ([qf.H(0), qf.RZ(0.2, 3), qf.CNOT(1, 2), qf.RY(0.4, 4)], 18)]
    for gl, per_amp in cases:
        # (diagonals kept as records of their own: folded into a 2x2 matrix
        # they turn it into a general one)
        plan = fusion.HostPlan(gl, tuple(range(n)), tile_bits=8, low_bits=3, block=16,
                               fuse_width=1, mode='reg', merge_records=False)
        assert plan.flops() == [per_amp * 2 ** n], [g.name for g in gl]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_virtualize_swaps_is_the_same_circuit(seed):
    """virtualize_swaps: the renamed gate list with its restoring swaps,
    applied gate by gate by the oracle, gives the original circuit's state;
    no SWAP-type gate is left before the restoring tail of plain swaps, at
    most n - 1 of them (one per displaced bit, less one per cycle)."""
    n = 9
    rng = np.random.default_rng(seed)
    names = [('SWAP', 2), ('ISWAP', 2), ('RX', 1), ('CNOT', 2), ('CZ', 2), ('H', 1),
             ('CCNOT', 3), ('PSWAP', 2)]
    desc = []
    for _ in range(60):
        nm, k = names[int(rng.integers(len(names)))]
        params = (float(rng.uniform(0, 6)),) if nm in ('RX', 'PSWAP') else ()
        desc.append((nm, params, tuple(int(q) for q in rng.permutation(n)[:k])))
    gates = build_gates(desc)
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    psi = (psi / np.linalg.norm(psi)).reshape([2] * n)
    ref = O.run_circuit(desc, qubits=tuple(range(n)), ket=psi)
    bits = [tuple(n - 1 - q for q in g.qubits) for g in gates]
    kinds, nbits, removed = fusion.virtualize_swaps([g.kind for g in gates], bits, n)
    assert removed == sum(d[0] in ('SWAP', 'ISWAP', 'PSWAP') for d in desc)
    ket = psi
    for kd, b in zip(kinds, nbits):
        m = kd.matrix.reshape([2] * (2 * len(b)))
        ket = O.tensormul(m, ket, [n - 1 - x for x in b])
    assert np.max(np.abs(ket - ref)) < 1e-12
    # the tail: plain swaps only; before it, no SWAP-type gate remains
    tail = 0
    while tail < len(kinds) and fusion.exchange_pattern(kinds[-1 - tail]) is not None \
            and fusion.exchange_pattern(kinds[-1 - tail])[0] == 'swap':
        tail += 1
    for kd in kinds[:len(kinds) - tail]:
        pat = fusion.exchange_pattern(kd) if kd.kind == 'monomial' else None
        assert pat is None or pat[0] != 'swap' or pat[3]
    assert tail <= n - 1

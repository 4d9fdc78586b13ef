"""Pins for the CPU oracle (oracle/qf_oracle.py).

Each oracle function is checked against something that does not restate it:
the reference's own known-answer tests (file:line cited), closed forms,
independent algorithms (transpose+matmul, the full Kronecker operator) and
brute force on small inputs.  A planted mistake (wrong axis order, a
transposed gate, a dropped global phase, a sign error in a gradient) fails
at least one of these.
"""
import numpy as np
import pytest
from math import pi, sqrt
from scipy.linalg import expm, fractional_matrix_power

from oracle import qf_oracle as O

RNG = np.random.default_rng(20240611)


def rand_state(n, rng=RNG):
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    v /= np.linalg.norm(v)
    return v.reshape([2] * n)


def rand_unitary(k, rng=RNG):
    a = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
    q, r = np.linalg.qr(a)
    return q * (np.diag(r) / np.abs(np.diag(r)))


# ---- tensormul: three independent formulations agree ----------------------

@pytest.mark.parametrize("n,k", [(1, 1), (3, 1), (4, 2), (5, 3), (6, 2), (7, 3), (8, 4)])
def test_tensormul_vs_matmul_and_full_operator(n, k):
    for trial in range(3):
        axes = list(RNG.permutation(n)[:k])
        u = rand_unitary(k)
        psi = rand_state(n)
        a = O.tensormul(u.reshape([2] * (2 * k)), psi, axes)
        b = O.tensormul_matmul(u.reshape([2] * (2 * k)), psi, axes)
        c = (O.full_operator(u, n, axes) @ psi.reshape(-1)).reshape([2] * n)
        assert np.max(np.abs(a - b)) < 1e-13
        assert np.max(np.abs(a - c)) < 1e-13


def test_full_operator_kron_closed_form():
    # gate on the two leading axes of a 4-qubit state = kron(U, I4)
    u = rand_unitary(2)
    op = O.full_operator(u, 4, [0, 1])
    assert np.allclose(op, np.kron(u, np.eye(4)), atol=1e-15)
    # gate on the last axis = kron(I8, U)
    u1 = rand_unitary(1)
    assert np.allclose(O.full_operator(u1, 4, [3]), np.kron(np.eye(8), u1))
    # reversed qubit order on a 2-qubit gate = SWAP U SWAP
    sw = O.gate_matrix('SWAP')
    assert np.allclose(O.full_operator(u, 2, [1, 0]), sw @ u @ sw)


# ---- index convention KATs ---------------------------------------------------

def test_bits_kat():
    # tests/test_gates.py:59-77
    x = O.gate_matrix('X')
    for q in range(4):
        ket = O.run_gate(O.zero_state(4), (0, 1, 2, 3), x, (q,))
        idx = [0, 0, 0, 0]
        idx[q] = 1
        assert ket[tuple(idx)] == 1.0
        assert np.argmax(np.abs(ket.reshape(-1))) == 1 << (3 - q)
    ket = O.zero_state(8)
    for q in (2, 4, 6):
        ket = O.run_gate(ket, tuple(range(8)), x, (q,))
    assert ket[0, 0, 1, 0, 1, 0, 1, 0] == 1.0


def test_state_labels_kat():
    # tests/test_states.py:54-66 (Quil labelling: qubits = 3,2,1,0)
    qubits = (3, 2, 1, 0)
    ket = O.zero_state(4)
    ket = O.run_gate(ket, qubits, O.gate_matrix('X'), (0,))
    ket = O.run_gate(ket, qubits, O.gate_matrix('X'), (1,))
    assert ket[0, 0, 1, 1] == 1.0


def test_cz_ccnot_cswap_kat():
    # tests/test_stdgates.py:85-130
    ket = O.run_gate(O.zero_state(2), (0, 1), O.gate_matrix('CZ'), (0, 1))
    assert ket[0, 0] == 1.0
    for q in (0, 1):
        ket = O.run_gate(ket, (0, 1), O.gate_matrix('X'), (q,))
    ket = O.run_gate(ket, (0, 1), O.gate_matrix('CZ'), (0, 1))
    assert -ket[1, 1] == 1.0

    qs = (0, 1, 2)
    ket = O.run_gate(O.zero_state(3), qs, O.gate_matrix('CCNOT'), qs)
    assert ket[0, 0, 0] == 1.0
    ket = O.run_gate(ket, qs, O.gate_matrix('X'), (1,))
    ket = O.run_gate(ket, qs, O.gate_matrix('CCNOT'), qs)
    assert ket[0, 1, 0] == 1.0
    ket = O.run_gate(ket, qs, O.gate_matrix('X'), (0,))
    ket = O.run_gate(ket, qs, O.gate_matrix('CCNOT'), qs)
    assert ket[1, 1, 1] == 1.0

    ket = O.run_gate(O.zero_state(3), qs, O.gate_matrix('X'), (1,))
    ket = O.run_gate(ket, qs, O.gate_matrix('CSWAP'), qs)
    assert ket[0, 1, 0] == 1.0
    ket = O.run_gate(ket, qs, O.gate_matrix('X'), (0,))
    ket = O.run_gate(ket, qs, O.gate_matrix('CSWAP'), qs)
    assert ket[1, 0, 1] == 1.0


# ---- readouts ---------------------------------------------------------------

def test_expectation_kats():
    # tests/test_states.py:133-140: H|0> with diag [0.4, 0.6] -> 0.5
    ket = O.run_gate(O.zero_state(1), (0,), O.gate_matrix('H'), (0,))
    assert abs(O.expectation(ket, [0.4, 0.6]) - 0.5) < 1e-15
    # tests/test_gates.py:150-164: 42 on |0000>, 2.5 on W(4)
    m = np.zeros([2] * 4)
    m[0, 0, 0, 0] = 42
    m[1, 0, 0, 0] = 1
    m[0, 1, 0, 0] = 2
    m[0, 0, 1, 0] = 3
    m[0, 0, 0, 1] = 4
    assert O.expectation(O.zero_state(4), m) == 42
    w = np.zeros(16, dtype=complex)
    for q in range(4):
        w[1 << (3 - q)] = 1 / sqrt(4)
    assert abs(O.expectation(w.reshape([2] * 4), m) - 2.5) < 1e-15


def test_probabilities_norm_inner_closed_forms():
    psi = rand_state(6)
    phi = rand_state(6)
    p = O.probabilities(psi)
    assert abs(p.sum() - 1.0) < 1e-14
    assert np.allclose(p.reshape(-1), (psi.conj() * psi).real.reshape(-1), atol=1e-16)
    assert abs(O.norm(psi) - 1.0) < 1e-14
    assert abs(O.norm(2 * psi) - 4.0) < 1e-13          # squared norm (qubits.py:180-182)
    s = O.inner(psi, phi)
    assert abs(s - np.sum(psi.conj() * phi)) < 1e-14
    assert abs(O.inner(phi, psi) - np.conj(s)) < 1e-14   # conjugate-symmetric


def test_expectation_z_closed_form():
    # product state RY(t_i)|0> on each qubit: <Z_i> = cos(t_i)
    n = 5
    ts = RNG.uniform(0, 2 * pi, size=n)
    ket = O.zero_state(n)
    for q in range(n):
        ket = O.run_gate(ket, tuple(range(n)), O.gate_matrix('RY', ts[q]), (q,))
    assert np.allclose(O.expectation_z(ket), np.cos(ts), atol=1e-14)


# ---- gate matrices ----------------------------------------------------------

ONEQ = ['X', 'Y', 'Z', 'H', 'S', 'T', 'S_H', 'T_H']


def test_gate_closed_forms():
    t = 1.2345
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0, -1.0]).astype(complex)
    assert np.allclose(O.gate_matrix('RX', t), expm(-0.5j * t * X))
    assert np.allclose(O.gate_matrix('RY', t), expm(-0.5j * t * Y))
    assert np.allclose(O.gate_matrix('RZ', t), expm(-0.5j * t * Z))
    assert np.allclose(O.gate_matrix('PHASE', t), np.diag([1, np.exp(1j * t)]))
    zz = np.kron(Z, Z)
    assert np.allclose(O.gate_matrix('ZZ', t), expm(-0.5j * pi * t * zz))
    xx = np.kron(X, X)
    assert np.allclose(O.gate_matrix('XX', t), expm(-0.5j * pi * t * xx))
    yy = np.kron(Y, Y)
    assert np.allclose(O.gate_matrix('YY', t), expm(-0.5j * pi * t * yy))
    # tests/test_stdgates.py:395-424: RN reduces to RX/RY/RZ
    assert np.allclose(O.gate_matrix('RN', t, 1, 0, 0), O.gate_matrix('RX', t))
    assert np.allclose(O.gate_matrix('RN', t, 0, 1, 0), O.gate_matrix('RY', t))
    assert np.allclose(O.gate_matrix('RN', t, 0, 0, 1), O.gate_matrix('RZ', t))
    # tests/test_stdgates.py:176-183
    assert np.allclose(O.gate_matrix('PSWAP', 0), O.gate_matrix('SWAP'))
    assert np.allclose(O.gate_matrix('PSWAP', pi / 2), O.gate_matrix('ISWAP'))
    assert np.allclose(O.gate_matrix('PISWAP', pi / 4), O.gate_matrix('ISWAP'))
    # TX/TY/TZ(t) = X^t etc: TX(1) = X
    assert np.allclose(O.gate_matrix('TX', 1.0), X)
    assert np.allclose(O.gate_matrix('TY', 1.0), Y)
    assert np.allclose(O.gate_matrix('TZ', 1.0), Z)
    assert np.allclose(O.gate_matrix('TH', 1.0), O.gate_matrix('H'))
    assert np.allclose(O.gate_matrix('T') @ O.gate_matrix('T'), O.gate_matrix('S'))
    # CAN(0,0,t) = ZZ(t); EXCH(t) = CAN(t,t,t)
    assert np.allclose(O.gate_matrix('CAN', 0, 0, t), O.gate_matrix('ZZ', t))


def test_gates_unitary():
    names = ONEQ + ['CZ', 'CNOT', 'SWAP', 'ISWAP', 'CCNOT', 'CSWAP']
    for nm in names:
        u = O.gate_matrix(nm)
        assert np.allclose(u @ u.conj().T, np.eye(u.shape[0]), atol=1e-15)
    for nm in ['PHASE', 'RX', 'RY', 'RZ', 'ZZ', 'XX', 'YY', 'PSWAP', 'PISWAP',
               'CPHASE', 'CPHASE00', 'CPHASE01', 'CPHASE10', 'TX', 'TY', 'TZ', 'TH', 'EXCH']:
        u = O.gate_matrix(nm, 0.777)
        assert np.allclose(u @ u.conj().T, np.eye(u.shape[0]), atol=1e-14)


# ---- circuits: the reference's golden wavefunctions --------------------------

def _fs_angle(a, b):
    # Fubini-Study angle (quantumflow/qubits.py:282-307)
    r = abs(np.vdot(a, b)) / sqrt(abs(np.vdot(a, a) * np.vdot(b, b)))
    return np.arccos(min(r, 1.0))


TRUE_KET = np.array([0.00167784 + 1.00210180e-05j, 0.50000000 - 4.99997185e-01j,
                     0.50000000 - 4.99997185e-01j, 0.00167784 + 1.00210180e-05j])


def qvm_circuit():
    # tests/test_circuits.py:43-60 (ReferenceQVM wavefunction)
    return [('RY', (pi / 2,), (0,)), ('RX', (pi,), (0,)), ('RY', (pi / 2,), (1,)),
            ('RX', (pi,), (1,)), ('CNOT', (), (0, 1)), ('RX', (-pi / 2,), (1,)),
            ('RY', (4.71572463191,), (1,)), ('RX', (pi / 2,), (1,)),
            ('CNOT', (), (0, 1)), ('RX', (-2 * 2.74973750579,), (0,)),
            ('RX', (-2 * 2.74973750579,), (1,))]


def test_qvm_wavefunction_kat():
    ket = O.run_circuit(qvm_circuit())
    assert _fs_angle(ket.reshape(-1), TRUE_KET) < 1e-6
    # 8 significant digits, equal up to a global phase
    ph = np.vdot(ket.reshape(-1), TRUE_KET)
    ph /= abs(ph)
    assert np.max(np.abs(ket.reshape(-1) * ph - TRUE_KET)) < 2e-7


def test_qvm_wavefunction_turns_kat():
    # tests/test_circuits.py:84-103 with TX/TY
    c = [('TY', (1 / 2,), (0,)), ('TX', (1,), (0,)), ('TY', (1 / 2,), (1,)),
         ('TX', (1,), (1,)), ('CNOT', (), (0, 1)), ('TX', (-1 / 2,), (1,)),
         ('TY', (4.71572463191 / pi,), (1,)), ('TX', (1 / 2,), (1,)),
         ('CNOT', (), (0, 1)), ('TX', (-2 * 2.74973750579 / pi,), (0,)),
         ('TX', (-2 * 2.74973750579 / pi,), (1,))]
    ket = O.run_circuit(c)
    assert _fs_angle(ket.reshape(-1), TRUE_KET) < 1e-6


def qft_gates(qubits):
    # quantumflow/circuits.py:170-185 (H, CPHASE ladder, then reversal SWAPs)
    n = len(qubits)
    g = []
    for n0 in range(n):
        g.append(('H', (), (qubits[n0],)))
        for n1 in range(n0 + 1, n):
            g.append(('CPHASE', (pi / 2 ** (n1 - n0),), (qubits[n1], qubits[n0])))
    for i in range(n // 2):
        g.append(('SWAP', (), (qubits[i], qubits[n - 1 - i])))
    return g


def test_qft_kat():
    # tests/test_circuits.py:190-203
    c = [('X', (), (2,))] + qft_gates([0, 1, 2])
    ket = O.run_circuit(c, qubits=(0, 1, 2))
    true = np.array([0.35355339 + 0.j, 0.25000000 + 0.25j, 0.00000000 + 0.35355339j,
                     -0.25000000 + 0.25j, -0.35355339 + 0.j, -0.25000000 - 0.25j,
                     0.00000000 - 0.35355339j, 0.25000000 - 0.25j])
    assert np.max(np.abs(ket.reshape(-1) - true)) < 1e-8


def adder_gates(addend0, addend1, carry):
    # quantumflow/circuits.py:306-356 (Cuccaro ripple-carry adder)
    def maj(q0, q1, q2):
        return [('CNOT', (), (q2, q1)), ('CNOT', (), (q2, q0)), ('CCNOT', (), (q0, q1, q2))]

    def uma(q0, q1, q2):
        return [('CCNOT', (), (q0, q1, q2)), ('CNOT', (), (q2, q0)), ('CNOT', (), (q0, q1))]
    qubits = [carry[0]] + [q for pair in zip(reversed(addend1), reversed(addend0))
                           for q in pair] + [carry[1]]
    g = []
    for i in range(0, len(qubits) - 3, 2):
        g += maj(*qubits[i:i + 3])
    g.append(('CNOT', (), (qubits[-2], qubits[-1])))
    for i in reversed(range(0, len(qubits) - 3, 2)):
        g += uma(*qubits[i:i + 3])
    return g


def test_adder_exact_kat():
    # tests/test_circuits.py:298-350: exact integer sums on basis states
    g = adder_gates([0, 1, 2], [3, 4, 5], [6, 7])
    from collections import Counter
    assert Counter(x[0] for x in g) == {'CNOT': 13, 'CCNOT': 6}   # :374-384
    for c0 in range(2):
        for a0 in range(8):
            for a1 in range(8):
                bits = [int(b) for b in format(a0, '03b')] + \
                       [int(b) for b in format(a1, '03b')] + [c0, 0]
                state = np.zeros([2] * 8)
                state[tuple(bits)] = 1
                ket = O.run_circuit(g, qubits=tuple(range(8)), ket=O.astensorproduct(state))
                flat = ket.reshape(-1)
                nz = np.nonzero(flat)[0]
                assert len(nz) == 1 and flat[nz[0]] == 1.0      # exact
                out = [(nz[0] >> (7 - q)) & 1 for q in range(8)]
                res = int(''.join(str(out[q]) for q in [7, 3, 4, 5]), 2)
                assert res == a0 + a1 + c0


W16_PAIRS = [(7, 8, 8), (4, 3, 3), (11, 12, 12), (2, 1, 1), (5, 6, 6), (10, 9, 9),
             (13, 14, 14), (1, 0, 0), (2, 3, 3), (5, 4, 4), (6, 7, 7), (9, 8, 8),
             (10, 11, 11), (13, 12, 12), (14, 15, 15)]


def w16_gates():
    # examples/state_prep_w16.py:24-111 (sqrt(ISWAP), S, Z and SWAP ladder)
    sq = fractional_matrix_power(O.gate_matrix('ISWAP'), 0.5)   # ops.py:175-184
    g = [('X', (), (7,))]
    swaps = {0: [(7, 6), (6, 5), (5, 4), (8, 9), (9, 10), (10, 11)],
             2: [(3, 2), (4, 5), (11, 10), (12, 13)]}
    for i, (a, b, s) in enumerate(W16_PAIRS):
        g.append((sq, (a, b)))
        g.append(('S', (), (s,)))
        g.append(('Z', (), (s,)))
        g += [('SWAP', (), p) for p in swaps.get(i, [])]
    return g


def test_w16_kat():
    # tests/test_examples.py:29-32: prepares W(16) up to a global phase
    ket = O.run_circuit(w16_gates(), qubits=tuple(range(16)))
    w = np.zeros(2 ** 16, dtype=complex)
    for q in range(16):
        w[1 << (15 - q)] = 1 / 4
    assert _fs_angle(ket.reshape(-1), w) < 1e-6


def test_circuit_unitary_matches_product():
    # circuits.py:111-118: asgate == product of full operators (last gate leftmost)
    c = qvm_circuit()
    u = O.circuit_unitary(c, (0, 1))
    ref = np.eye(4, dtype=complex)
    for name, params, qs in c:
        ref = O.full_operator(O.gate_matrix(name, *params), 2, list(qs)) @ ref
    assert np.allclose(u, ref, atol=1e-14)


# ---- DAG --------------------------------------------------------------------

def test_dag_depth_kat():
    # tests/test_dagcircuit.py:97-106
    q = [g[-1] for g in qft_gates([0, 1, 2, 3])]
    assert O.dag_depth(q) == 8
    ghz = [(0,)] + [(i, i + 1) for i in range(4)]     # circuits.py:359-371
    assert O.dag_depth(ghz) == 5
    assert O.dag_depth(ghz, local=False) == 4
    assert len(O.dag_layers(ghz)) == 5


# ---- QAOA -------------------------------------------------------------------

def test_graph_cuts_kat():
    # tests/test_qaoa.py:40-55
    cut = O.graph_cuts(3, [(0, 1), (1, 2)])
    assert np.allclose(cut.reshape(-1), [0.0, 1.0, 2.0, 1.0, 1.0, 2.0, 1.0, 0.0])
    cut = O.graph_cuts(3, [(0, 1), (1, 2), (0, 2)], [0.1, 0.4, 0.35])
    assert np.allclose(cut.reshape(-1), [0., 0.75, 0.5, 0.45, 0.45, 0.5, 0.75, 0.])


def test_graph_cuts_vs_loop_bitwise():
    for n in (3, 5, 7):
        edges = [(i, j) for i in range(n) for j in range(i + 1, n) if RNG.random() < 0.5]
        w = list(RNG.uniform(0.1, 2.0, size=len(edges)))
        assert np.array_equal(O.graph_cuts(n, edges), O.graph_cuts_loop(n, edges))
        assert np.array_equal(O.graph_cuts(n, edges, w), O.graph_cuts_loop(n, edges, w))


def test_maxcut_bruteforce_small():
    # ring of 8: max cut 8; triangle: 2; star K1,3: 3
    assert O.maxcut_bruteforce(8, [(i, (i + 1) % 8) for i in range(8)]) == 8
    assert O.maxcut_bruteforce(3, [(0, 1), (1, 2), (0, 2)]) == 2
    assert O.maxcut_bruteforce(4, [(0, 1), (1, 2), (1, 3)]) == 3


def test_qubo_circuit_counts_kat():
    # tests/test_qaoa.py:17-37
    g = O.qubo_circuit([0, 1, 2], [(0, 1), (1, 2)], 2, [1, 1], [2, 2], [0.1, 0.4])
    assert len(g) == 13
    g = O.qubo_circuit([0, 1, 2], [(0, 1), (1, 2)], 2, [1, 1], [2, 2], [0.1, 0.4],
                       node_weights={0: 4})
    assert len(g) == 15


def _p1_analytic(n, edges, beta, gamma):
    """Closed form of p=1 QAOA MaxCut on an unweighted graph (Wang, Hadfield,
    Jiang, Rieffel, PRA 97, 022304 (2018), Thm 1) with the reference's
    parameterisation: cost exp(-i gamma C) up to a global phase and mixer
    RX(beta) = exp(-i (beta/2) X), i.e. the paper's beta is beta/2."""
    b = beta / 2.0
    g = gamma
    adj = [set() for _ in range(n)]
    for u, v in edges:
        adj[u].add(v)
        adj[v].add(u)
    tot = 0.0
    for u, v in edges:
        du, dv = len(adj[u]) - 1, len(adj[v]) - 1
        lam = len(adj[u] & adj[v])
        term = 0.5 + 0.25 * np.sin(4 * b) * np.sin(g) * (np.cos(g) ** du + np.cos(g) ** dv)
        term -= 0.25 * np.sin(2 * b) ** 2 * np.cos(g) ** (du + dv - 2 * lam) * \
            (1 - np.cos(2 * g) ** lam)
        tot += term
    return tot


def test_qaoa_p1_closed_form():
    for trial in range(4):
        n = 6
        edges = [(i, j) for i in range(n) for j in range(i + 1, n) if RNG.random() < 0.5]
        beta, gamma = RNG.uniform(0, pi), RNG.uniform(0, pi)
        e = O.qaoa_expectation(n, edges, [beta], [gamma])
        assert abs(e - _p1_analytic(n, edges, beta, gamma)) < 1e-12


def test_qaoa_fast_vs_gate_by_gate():
    n = 7
    edges = [(i, j) for i in range(n) for j in range(i + 1, n) if RNG.random() < 0.45]
    beta = RNG.uniform(0, pi / 2, size=3)
    gamma = RNG.uniform(0, pi, size=3)
    a = O.qaoa_state(n, edges, beta, gamma)
    b = O.qaoa_state_fast(n, edges, beta, gamma)
    assert np.max(np.abs(a - b)) < 1e-13        # includes the e^{i gamma W/2} phase


def test_qaoa_gradients_agree():
    n = 5
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)]
    beta = RNG.uniform(0, pi / 2, size=2)
    gamma = RNG.uniform(0, pi, size=2)
    sb, sg = O.qaoa_grad_shift(n, edges, beta, gamma)
    fb, fg = O.qaoa_grad_fd(n, edges, beta, gamma, h=1e-3, order=4)
    assert np.max(np.abs(sb - fb)) < 1e-9
    assert np.max(np.abs(sg - fg)) < 1e-9
    fb2, fg2 = O.qaoa_grad_fd(n, edges, beta, gamma, h=1e-6, order=2, fast=False)
    assert np.max(np.abs(sb - fb2)) < 1e-7
    assert np.max(np.abs(sg - fg2)) < 1e-7


def test_qaoa_torch_autograd_vs_shift_rule():
    # the torch restatement: value = the gate-by-gate oracle, autograd
    # gradient = the exact per-occurrence parameter-shift rule
    import torch
    n = 6
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 0), (0, 3)]
    beta = RNG.uniform(0, pi / 2, size=3)
    gamma = RNG.uniform(0, pi, size=3)
    b = torch.tensor(beta, requires_grad=True)
    g = torch.tensor(gamma, requires_grad=True)
    v = O.qaoa_value_torch(n, edges, b, g)
    v.backward()
    assert abs(v.item() - O.qaoa_expectation(n, edges, beta, gamma)) < 1e-12
    sb, sg = O.qaoa_grad_shift(n, edges, beta, gamma)
    assert np.max(np.abs(b.grad.numpy() - sb)) < 1e-12
    assert np.max(np.abs(g.grad.numpy() - sg)) < 1e-12


def test_qaoa_fast_with_given_cuts():
    n = 8
    edges = [(i, (i * 3 + 1) % n) for i in range(n) if i != (i * 3 + 1) % n]
    beta, gamma = RNG.uniform(0, 1, size=2), RNG.uniform(0, 1, size=2)
    cuts = O.graph_cuts(n, edges)
    assert O.qaoa_expectation_fast(n, edges, beta, gamma, cuts=cuts) == \
        O.qaoa_expectation_fast(n, edges, beta, gamma)


def test_linear_ramp():
    b, g = O.linear_ramp(4, 0.8)
    assert np.allclose(b + g, 0.8)
    assert np.all(np.diff(g) > 0) and np.all(np.diff(b) < 0)


def test_maxcut_star_ratio_kat():
    # tests/test_qaoa_maxcut.py:12-21: 4-node star reaches ratio > 0.95 at
    # p=5 with the reference's init N(0.5, 0.01) and lr 0.01 (TF-only there)
    rng = np.random.default_rng(7)
    b0 = rng.normal(0.5, 0.01, size=5)
    g0 = rng.normal(0.5, 0.01, size=5)
    ratio, b, g = O.maxcut_qaoa_gd(4, [(0, 1), (1, 2), (1, 3)], steps=5,
                                   beta0=b0, gamma0=g0, max_opt_steps=4000)
    assert ratio > 0.95

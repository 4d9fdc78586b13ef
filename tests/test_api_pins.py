"""Host-side pins of API pieces whose values the GPU path consumes (no GPU).

* Every class of STDGATES: its matrix bit for bit against the oracle's
  restatement of the reference class, its ``.H`` as the conjugate
  transpose, and its ``**t`` by (g**1/2)(g**1/2) = g and g**-1 = g.H.
* DAGCircuit depth, layers, components, next/prev against the oracle's
  wire-DAG restatement and the reference's known answers
  (tests/test_dagcircuit.py:97-160).
* Channel host algebra (trace, partial_trace, chi, H, permute) against
  closed forms: tr(U (x) conj U) = |tr U|^2, and the four-index partial
  trace of a product channel U1 (x) U2 over the first qubit =
  (sum_a |U1_aa|^2) times the channel of U2.
"""
import numpy as np
import pytest

from oracle import qf_oracle as O


@pytest.fixture(scope="module")
def qf():
    import quantumflow_amd as qf
    return qf


def _instances(qf, rng):
    for name, cls in sorted(qf.STDGATES.items()):
        pn = getattr(cls, '_pnames', ())
        for _ in range(3):
            params = [float(rng.uniform(-2 * np.pi, 2 * np.pi)) for _ in pn]
            if name == 'RN':
                # RN takes a unit rotation axis (stdgates.py:604-636)
                axis = rng.normal(size=3)
                params[1:] = [float(a) for a in axis / np.linalg.norm(axis)]
            yield name, cls, params


def test_stdgates_bitwise_vs_oracle(qf):
    rng = np.random.default_rng(31)
    assert len(qf.STDGATES) == 36
    seen = set()
    for name, cls, params in _instances(qf, rng):
        g = cls(*params)
        # _TPow classes reduce t mod 2 before building the matrix, as the
        # reference does (stdgates.py:685,717,748)
        used = list(g.params.values()) if g.params else params
        ref = O.gate_matrix(name, *used)
        assert g.matrix.dtype == np.complex128
        assert np.array_equal(g.matrix, ref), name
        seen.add(name)
    assert len(seen) == 36


def test_stdgates_hermitian_conjugate_and_powers(qf):
    rng = np.random.default_rng(32)
    for name, cls, params in _instances(qf, rng):
        g = cls(*params)
        m = g.matrix
        h = g.H
        assert np.allclose(h.matrix, m.conj().T, atol=1e-14, rtol=0), name
        assert tuple(h.qubits) == tuple(g.qubits)
        half = g ** 0.5
        assert np.allclose(half.matrix @ half.matrix, m, atol=1e-12, rtol=0), name
        inv = g ** -1
        assert np.allclose(inv.matrix, m.conj().T, atol=1e-12, rtol=0), name


def test_stdgates_identities(qf):
    # reference tests/test_stdgates.py:70-72, 176-183, 395-424
    assert np.allclose(qf.RX(np.pi).matrix, -1j * qf.X().matrix)
    assert np.array_equal(qf.PSWAP(0).matrix, qf.SWAP().matrix)
    t = 0.37
    assert np.array_equal(qf.RN(t, 1, 0, 0).matrix, qf.RX(t).matrix)
    assert np.array_equal(qf.TX(1).matrix, O.gate_matrix('TX', 1))
    assert np.allclose(qf.TX(1).matrix, qf.X().matrix, atol=1e-15)


def _qubits_of(circ):
    return [tuple(e.qubits) for e in circ]


def test_dag_depth_layers_kats(qf):
    circ = qf.qft_circuit([0, 1, 2, 3])
    dag = qf.DAGCircuit(circ)
    assert dag.depth() == 8 == O.dag_depth(_qubits_of(circ))
    circ = qf.ghz_circuit(range(5))
    dag = qf.DAGCircuit(circ)
    assert dag.depth() == 5 == O.dag_depth(_qubits_of(circ))
    assert dag.depth(local=False) == 4 == O.dag_depth(_qubits_of(circ), local=False)
    layers = dag.layers()
    assert len(layers.elements) == dag.depth()
    assert dag.size() == len(circ.elements)


def test_dag_layers_vs_oracle_random(qf):
    rng = np.random.default_rng(33)
    one = [qf.H, qf.X, qf.T]
    two = [qf.CNOT, qf.CZ, qf.SWAP]
    for trial in range(20):
        nq = int(rng.integers(2, 8))
        gates = []
        for _ in range(int(rng.integers(1, 40))):
            if rng.uniform() < 0.5:
                gates.append(one[rng.integers(3)](int(rng.integers(nq))))
            else:
                a, b = rng.choice(nq, size=2, replace=False)
                gates.append(two[rng.integers(3)](int(a), int(b)))
        circ = qf.Circuit(gates)
        dag = qf.DAGCircuit(circ)
        qs = _qubits_of(circ)
        assert dag.depth() == O.dag_depth(qs)
        assert dag.depth(local=False) == O.dag_depth(qs, local=False)
        want = O.dag_layers(qs)
        got = dag.layers()
        assert len(got.elements) == len(want)
        for layer, idx in zip(got.elements, want):
            assert [id(e) for e in layer.elements] == [id(gates[i]) for i in idx]
        # every element is yielded once, in circuit order
        assert [id(e) for e in dag] == [id(g) for g in gates]


def test_dag_components_kats(qf):
    # reference tests/test_dagcircuit.py:116-145
    circ = qf.Circuit()
    circ += qf.H(0)
    circ += qf.H(1)
    assert qf.DAGCircuit(circ).component_nb() == 2
    circ += qf.CNOT(0, 1)
    assert qf.DAGCircuit(circ).component_nb() == 1
    circ = qf.Circuit()
    circ.extend(qf.ghz_circuit([0, 2, 4, 6, 8]))
    circ.extend(qf.ghz_circuit([1, 3, 5, 7, 9]))
    dag = qf.DAGCircuit(circ)
    assert dag.component_nb() == 2
    comps = dag.components()
    assert sorted(tuple(c.qubits) for c in comps) == [(0, 2, 4, 6, 8), (1, 3, 5, 7, 9)]
    assert sum(c.size() for c in comps) == dag.size()
    circ.extend(qf.qft_circuit([0, 2, 4, 6]))
    circ.extend(qf.ghz_circuit([1, 3, 5, 7]))
    circ += qf.H(10)
    dag = qf.DAGCircuit(circ)
    comps = dag.components()
    assert dag.component_nb() == 3 and len(comps) == 3


def test_dag_next_prev_kats(qf):
    # reference tests/test_dagcircuit.py:148-160
    circ = qf.ghz_circuit([0, 2, 4, 6, 8])
    elem = circ.elements[3]
    dag = qf.DAGCircuit(circ)
    assert dag.next_element(elem, elem.qubits[1]) is circ.elements[4]
    assert dag.prev_element(elem, elem.qubits[0]) is circ.elements[2]
    assert dag.next_element(elem, elem.qubits[0]) == ('out', 4)
    assert dag.prev_element(elem, elem.qubits[1]) == ('in', 6)


def _unitary(rng, d):
    z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def test_channel_trace_and_partial_trace_closed_forms(qf):
    rng = np.random.default_rng(34)
    u1, u2 = _unitary(rng, 2), _unitary(rng, 2)
    u = np.kron(u1, u2)
    chan = qf.Gate(u, (0, 1)).aschannel()
    assert abs(chan.trace() - abs(np.trace(u)) ** 2) <= 1e-12
    pt = chan.partial_trace([0])
    assert pt.qubits == (1,)
    want = np.sum(np.abs(np.diag(u1)) ** 2) * np.kron(u2, u2.conj())
    assert np.allclose(pt.asoperator(), want, atol=1e-12, rtol=0)
    pt1 = chan.partial_trace([1])
    want1 = np.sum(np.abs(np.diag(u2)) ** 2) * np.kron(u1, u1.conj())
    assert np.allclose(pt1.asoperator(), want1, atol=1e-12, rtol=0)
    with pytest.raises(ValueError):
        chan.partial_trace([0, 1])


def test_channel_chi_h_permute(qf):
    rng = np.random.default_rng(35)
    u = _unitary(rng, 4)
    chan = qf.Gate(u, ('a', 'b')).aschannel()
    assert np.array_equal(chan.chi(), chan.choi())
    assert np.array_equal(chan.H.asoperator(), chan.asoperator().conj().T)
    # permuting the labels of U (x) conj U is the channel of the permuted U
    perm = chan.permute(('b', 'a'))
    assert perm.qubits == ('b', 'a')
    sw = O.gate_matrix('SWAP')
    up = sw @ u @ sw
    assert np.allclose(perm.asoperator(), np.kron(up, up.conj()), atol=1e-14, rtol=0)


def test_program_quil_text(qf):
    # reference tests/test_programs.py:95-99, 153-157 and the DECLARE form
    prog = qf.Program([qf.Call('BELL', params=[], qubits=[])])
    assert str(prog) == 'BELL\n'
    assert qf.Include('somefile.quil', qf.Program()).quil() == 'INCLUDE "somefile.quil"'
    assert qf.Declare('ro', 'BIT', 2).quil() == 'DECLARE ro BIT [2]'
    assert qf.Call('RX', [0.5], [3]).quil() == 'RX(0.5) 3'
    assert qf.JumpWhen('a', qf.Register()[0]).quil() == 'JUMP-WHEN @a ro[0]'
    assert qf.Program([qf.Label('x'), qf.Jump('x')]).quil() == 'LABEL @x\nJUMP @x\n'
    r = qf.Register('r', 'INT')
    assert r[1] == qf.Register('r', 'INT')[1] and hash(r[1]) == hash(qf.Register('r', 'INT')[1])
    assert r[0] < r[1] and str(r[1]) == 'r[1]'
    with pytest.raises(AssertionError):
        qf.Register('x', 'FLOAT')

"""Quil-style programs over the device state: the program counter, labels,
jumps, halting, classical memory and named-gate calls around the GPU gate,
measurement and reset kernels, against the reference's known answers
(tests/test_programs.py) and the oracle."""
import numpy as np
import pytest

from oracle import qf_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qf():
    import quantumflow_amd as qf
    assert qf.backend.gpu_available()
    return qf


def test_empty_and_nop(qf):
    ket = qf.Program().run()
    assert ket.qubits == () and ket.qubit_nb == 0
    prog = qf.Program([qf.Nop()])
    prog.run()
    prog.evolve()
    assert qf.Nop().qubits == () and qf.Nop().qubit_nb == 0


def test_labels_and_jumps(qf):
    from quantumflow_amd.programs import TARGETS, PC
    prog = qf.Program([qf.Label('Here'), qf.Nop(), qf.Label('There')])
    assert prog.run().memory[TARGETS] == {'Here': 0, 'There': 2}
    ro = qf.Register()
    prog = qf.Program()
    prog += qf.Move(ro[0], 0)
    prog += qf.Jump('There')
    prog += qf.Not(ro[0])
    prog += qf.Label('There')
    prog += qf.Not(ro[0])
    assert prog.run().memory[ro[0]] == 1
    prog += qf.JumpWhen('There', ro[0])
    assert prog.run().memory[ro[0]] == 0
    prog += qf.Not(ro[0])
    prog += qf.JumpUnless('There', ro[0])
    assert prog.run().memory[ro[0]] == 1
    halt = qf.Program([qf.Move(ro[0], 0), qf.Halt(), qf.Not(ro[0])])
    ket = halt.run()
    assert ket.memory[PC] == -1 and ket.memory[ro[0]] == 0
    wait = qf.Program([qf.Move(ro[0], 0), qf.Wait(), qf.Not(ro[0]), qf.Wait(), qf.Not(ro[0]),
                       qf.Wait(), qf.Not(ro[0])])
    assert wait.run().memory[ro[0]] == 1


def test_reset_and_measure(qf):
    ro = qf.Register()
    prog = qf.Program([qf.Move(ro[0], 1), qf.Call('X', params=[], qubits=[0]), qf.Reset(),
                       qf.Measure(0, ro[1])])
    ket = prog.run()
    assert ket.qubits == (0,)
    assert ket.memory[ro[0]] == 1 and ket.memory[ro[1]] == 0
    prog = qf.Program([qf.Call('X', [], [0]), qf.Call('X', [], [1]), qf.Reset(0),
                       qf.Measure(0, ('b', 0)), qf.Measure(1, ('b', 1))])
    ket = prog.run()
    assert ket.memory[('b', 0)] == 0 and ket.memory[('b', 1)] == 1


def test_measure_until(qf):
    # repeat X, H, measure until the outcome is 1 (a loop whose length the
    # draws decide)
    np.random.seed(7)
    prog = qf.Program([qf.Move(('c', 2), 1), qf.Label('redo'), qf.Call('X', [], [0]),
                       qf.Call('H', [], [0]), qf.Measure(0, ('c', 2)),
                       qf.JumpUnless('redo', ('c', 2))])
    assert prog.run().memory[('c', 2)] == 1


def test_bell_occupation_and_qvm_kat(qf):
    ket = qf.Program([qf.Call('H', [], [0]), qf.Call('CNOT', [], [0, 1])]).run()
    assert qf.states_close(ket, qf.ghz_state(2))
    prog = qf.Program([qf.Call('X', [], [0]), qf.Call('X', [], [1]), qf.Call('I', [], [2]),
                       qf.Call('I', [], [3])])
    ket = prog.run()
    assert ket.qubits == (0, 1, 2, 3)
    probs = qf.asarray(ket.probabilities())
    assert probs[1, 1, 0, 0] == 1.0 and probs[1, 1, 0, 1] == 0.0
    # the reference QVM's 2-qubit wavefunction (tests/test_programs.py:190-211)
    wf_true = np.array([0.00167784 + 1.00210180e-05j, 0.50000000 - 4.99997185e-01j,
                        0.50000000 - 4.99997185e-01j, 0.00167784 + 1.00210180e-05j])
    prog = qf.Program()
    prog += qf.Call('RY', [np.pi / 2], [0])
    prog += qf.Call('RX', [np.pi], [0])
    prog += qf.Call('RY', [np.pi / 2], [1])
    prog += qf.Call('RX', [np.pi], [1])
    prog += qf.Call('CNOT', [], [0, 1])
    prog += qf.Call('RX', [-np.pi / 2], [1])
    prog += qf.Call('RY', [4.71572463191], [1])
    prog += qf.Call('RX', [np.pi / 2], [1])
    prog += qf.Call('CNOT', [], [0, 1])
    prog += qf.Call('RX', [-2 * 2.74973750579], [0])
    prog += qf.Call('RX', [-2 * 2.74973750579], [1])
    assert qf.states_close(prog.run(), qf.State(wf_true.reshape(2, 2)))
    with pytest.raises(RuntimeError):
        qf.Program([qf.Call('BELL', [], [])]).run()


def test_program_matches_oracle_on_a_wide_state(qf):
    """A 16-qubit program of named-gate calls inside a classical counter loop
    (the body runs three times): the final state equals the oracle's run of
    the unrolled gate list."""
    rng = np.random.default_rng(9)
    n = 16
    body = []
    for _ in range(12):
        if rng.uniform() < 0.5:
            q = int(rng.integers(n))
            body.append(('RX', (float(rng.uniform(0, 6)),), (q,)))
        else:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            body.append(('CNOT', (), (a, b)))
    i = qf.Register('i', 'INT')
    k = qf.Register('k', 'INT')
    done = qf.Register('done', 'BIT')
    prog = qf.Program([qf.Declare('i', 'INT', 1), qf.Declare('done', 'BIT', 1),
                       qf.Move(k[0], 3), qf.Call('H', [], [0]), qf.Label('loop')])
    for name, params, qs in body:
        prog += qf.Call(name, list(params), list(qs))
    prog += qf.Add(i[0], 1)
    prog += qf.EQ(done[0], i[0], k[0])
    prog += qf.JumpUnless('loop', done[0])
    ket = prog.run(qf.zero_state(n))
    assert ket.memory[i[0]] == 3
    desc = [('H', (), (0,))] + body * 3
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    assert np.max(np.abs(ket.vec.asarray() - ref)) <= 1e-12


def test_classical_arithmetic(qf):
    r = qf.Register('r', 'REAL')
    prog = qf.Program([qf.Move(r[0], 6), qf.Move(r[1], 4), qf.Add(r[0], r[1]), qf.Sub(r[0], 1),
                       qf.Mul(r[0], 2), qf.Div(r[0], 3), qf.Neg(r[0]), qf.Exchange(r[0], r[1]),
                       qf.Move(r[2], 1), qf.And(r[2], 1), qf.Ior(r[2], 0), qf.Xor(r[2], 1),
                       qf.LT(r[3], r[1], r[0]), qf.GT(r[4], r[1], r[0]), qf.GE(r[5], r[0], r[0]),
                       qf.LE(r[6], r[1], r[0]), qf.EQ(r[7], r[0], r[0]),
                       qf.NE(r[8], r[0], r[1]), qf.NE(r[9], r[0], r[0])])
    m = prog.run().memory
    assert m[r[1]] == -6.0 and m[r[0]] == 4
    assert m[r[2]] == 0
    assert m[r[3]] and not m[r[4]] and m[r[5]] and m[r[6]] and m[r[7]]
    assert m[r[8]] and not m[r[9]]


def test_program_evolve_density(qf):
    """The same program on a density: X then a measurement into memory, and
    the density it leaves."""
    ro = qf.Register()
    prog = qf.Program([qf.Call('X', [], [0]), qf.Call('H', [], [1]), qf.Measure(0, ro[0])])
    rho = prog.evolve()
    assert rho.memory[ro[0]] == 1
    m = rho.asoperator().cpu().numpy()
    ref = O.density_from_state(O.run_circuit([('X', (), (0,)), ('H', (), (1,))],
                                             qubits=(0, 1)))
    assert np.max(np.abs(m - ref.reshape(4, 4))) <= 1e-12


# ---------------------------------------------------------------------------
# Quil text through quantumflow_amd.forest (tests/test_quil.py,
# tests/test_forest.py), on the device state

def _quil(name):
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'quil', name)
    with open(path) as f:
        return f.read()


def test_run_quil_files(qf):
    """tests/test_quil.py:36-44: the runnable fixtures run, with the memory
    and state their instructions fix."""
    P = qf.forest.quil_to_program
    ro = qf.Register()
    assert P(_quil('empty.quil')).run().qubit_nb == 0
    ket = P(_quil('hello_world.quil')).run()
    assert ket.memory[ro[0]] == 1
    ket = P(_quil('measure.quil')).run()
    assert ket.memory[ro[0]] == 1
    ket = P(_quil('classical_logic.quil')).run()
    m = ket.memory
    assert [m[ro[i]] for i in range(5)] == [0, 0, 0, 1, 0]
    ket = P(_quil('control_flow.quil')).run()
    assert ket.memory[ro[0]] == 0
    ket = P(_quil('bell.quil')).run()
    assert qf.states_close(ket, qf.ghz_state(2))
    # the reference QVM's wavefunction for this program
    # (tests/test_programs.py:190-211)
    wf_true = np.array([0.00167784 + 1.00210180e-05j, 0.50000000 - 4.99997185e-01j,
                        0.50000000 - 4.99997185e-01j, 0.00167784 + 1.00210180e-05j])
    ket = P(_quil('qaoa.quil')).run()
    assert qf.states_close(ket, qf.State(wf_true.reshape(2, 2)))


def test_qvm(qf):
    """The QVM status machine around the device state
    (forest/__init__.py:369-444)."""
    qvm = qf.forest.QuantumFlowQVM()
    qvm.load(_quil('bell.quil')).run().wait()
    assert qvm.status == 'done'
    assert qf.states_close(qvm.wavefunction(), qf.ghz_state(2))
    qvm.load('X 0\nMEASURE 0 ro[0]\nMEASURE 1 ro[1]').run().wait()
    assert qvm.read_from_memory_region(region_name='ro') == [1, 0]
    with pytest.raises(NotImplementedError):
        qvm.read_from_memory_region(region_name='ro', offsets=[0])
    prog = qf.Program([qf.Call('X', [], [2]), qf.Measure(2, ('c', 0))])
    qvm.load(prog).run().wait()
    assert qvm.wavefunction().memory[('c', 0)] == 1


def test_quil_round_trip_wide_state(qf):
    """A 14-qubit program of standard-gate calls, printed as Quil and parsed
    back, runs to the same state bit for bit, and to the oracle's state."""
    rng = np.random.default_rng(21)
    n = 14
    desc = [('H', (), (q,)) for q in range(n)]
    for _ in range(60):
        r = rng.uniform()
        if r < 0.5:
            name = ('RX', 'RY', 'RZ', 'PHASE')[int(rng.integers(4))]
            desc.append((name, (float(rng.uniform(-7, 7)),), (int(rng.integers(n)),)))
        else:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            desc.append((('CNOT', 'CZ', 'SWAP', 'ISWAP')[int(rng.integers(4))], (), (a, b)))
    prog = qf.Program([qf.Call(name, list(p), list(qs)) for name, p, qs in desc])
    parsed = qf.forest.quil_to_program(prog.quil())
    k0 = prog.run(qf.zero_state(n))
    k1 = parsed.run(qf.zero_state(n))
    a0, a1 = k0.vec.asarray(), k1.vec.asarray()
    assert np.array_equal(a0, a1)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    assert np.max(np.abs(a1 - ref)) <= 1e-12

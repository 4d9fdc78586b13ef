"""GPU parity of density matrices and Kraus channels against the oracle
(complex128, tolerance 1e-12 on matrix entries)."""
import numpy as np
import pytest

from oracle import qf_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qf():
    import quantumflow_amd as qf
    assert qf.backend.gpu_available()
    return qf


def _mat(rho):
    n = rho.qubit_nb
    return np.asarray(rho.asoperator().cpu().numpy()).reshape(2 ** n, 2 ** n)


def _rand_density(n, rng):
    g = rng.normal(size=(2 ** n, 2 ** n)) + 1j * rng.normal(size=(2 ** n, 2 ** n))
    m = g @ g.conj().T
    return m / np.trace(m)


def test_asdensity_trace_probabilities(qf):
    rng = np.random.default_rng(1)
    n = 6
    psi = rng.normal(size=[2] * n) + 1j * rng.normal(size=[2] * n)
    psi /= np.linalg.norm(psi)
    print("asdensity n", n, flush=True)
    rho = qf.State(psi).asdensity()
    assert np.max(np.abs(_mat(rho) - O.density_from_state(psi))) <= 1e-12
    assert abs(complex(rho.trace().item()) - 1.0) <= 1e-12
    probs = rho.probabilities().cpu().numpy().reshape(-1)
    assert np.max(np.abs(probs - np.abs(psi.reshape(-1)) ** 2)) <= 1e-12
    mixed = qf.mixed_density(4)
    assert np.max(np.abs(_mat(mixed) - np.eye(16) / 16)) == 0.0
    half = qf.Density(_mat(rho).reshape([2] * (2 * n)) * 2.0)
    assert np.max(np.abs(_mat(half.normalize()) - _mat(rho))) <= 1e-12


@pytest.mark.parametrize("name,params,qubits", [
    ("H", (), [3]), ("RX", (0.7,), [0]), ("CNOT", (), [4, 1]), ("ISWAP", (), [2, 5]),
    ("CZ", (), [0, 5]), ("CCNOT", (), [5, 0, 3]), ("PHASE", (1.1,), [2])])
def test_gate_evolve_matches_oracle(qf, name, params, qubits):
    rng = np.random.default_rng(len(name) + len(qubits))
    n = 6
    m0 = _rand_density(n, rng)
    print("evolve", name, qubits, flush=True)
    rho = qf.Density(m0.reshape([2] * (2 * n)))
    gate = getattr(qf, name)(*params, *qubits)
    out = gate.evolve(rho)
    ref = O.density_evolve(m0, O.gate_matrix(name, *params), n, qubits)
    assert np.max(np.abs(_mat(out) - ref)) <= 1e-12


def test_circuit_evolve_equals_pure_state_run(qf):
    rng = np.random.default_rng(7)
    n = 5
    circ = qf.Circuit()
    for _ in range(12):
        a, b = rng.choice(n, size=2, replace=False)
        circ += qf.RY(float(rng.uniform(0, 6)), int(a))
        circ += qf.CNOT(int(a), int(b))
    print("circuit evolve", flush=True)
    rho = circ.evolve(qf.zero_state(n).asdensity())
    ket = circ.run(qf.zero_state(n))
    psi = ket.tensor.cpu().numpy()
    assert np.max(np.abs(_mat(rho) - O.density_from_state(psi))) <= 1e-12


def test_kraus_channels_match_oracle(qf):
    rng = np.random.default_rng(9)
    n = 4
    m0 = _rand_density(n, rng)
    rho = qf.Density(m0.reshape([2] * (2 * n)))
    I2, X = np.eye(2), np.array([[0, 1], [1, 0]])
    Y, Z = np.array([[0, -1j], [1j, 0]]), np.diag([1.0, -1.0])
    p = 0.25
    cases = [
        (qf.Damping(p, 2), [(np.array([[1, 0], [0, np.sqrt(1 - p)]]), [2], 1.0),
                            (np.array([[0, np.sqrt(p)], [0, 0]]), [2], 1.0)]),
        (qf.Depolarizing(p, 0), [(I2, [0], 1 - p), (X, [0], p / 3), (Y, [0], p / 3),
                                 (Z, [0], p / 3)]),
        (qf.Dephasing(p, 3), [(I2, [3], 1 - p / 2), (Z, [3], p / 2)]),
    ]
    for chan, ops in cases:
        print("channel", type(chan).__name__, flush=True)
        out = chan.evolve(rho)
        ref = O.kraus_evolve(m0, ops, n)
        assert np.max(np.abs(_mat(out) - ref)) <= 1e-12
        assert abs(complex(out.trace().item()) - 1.0) <= 1e-12
    # closed form: damping sends |1><1| to diag(p, 1 - p)
    one = qf.State(np.array([0.0, 1.0])).asdensity()
    got = _mat(qf.Damping(p, 0).evolve(one))
    assert np.max(np.abs(got - np.diag([p, 1 - p]))) <= 1e-15


def test_two_qubit_kraus_channel_matches_oracle(qf):
    """A channel on two qubits (its superoperator is a dense 4-qubit gate,
    which no single kernel takes): operator evolutions summed by weights."""
    rng = np.random.default_rng(11)
    n = 5
    m0 = _rand_density(n, rng)
    rho = qf.Density(m0.reshape([2] * (2 * n)))
    chan = qf.Kraus([qf.RX(0.4, 3), qf.CNOT(3, 1), qf.H(1)], [0.5, 0.3, 0.2])
    out = chan.evolve(rho)
    ops = [(O.gate_matrix('RX', 0.4), [3], 0.5), (O.gate_matrix('CNOT'), [3, 1], 0.3),
           (O.gate_matrix('H'), [1], 0.2)]
    ref = O.kraus_evolve(m0, ops, n)
    assert np.max(np.abs(_mat(out) - ref)) <= 1e-12


def test_density_entry_points_reject_bad_sizes(qf):
    import torch
    from quantumflow_amd import _native
    lib = _native.load()
    buf = torch.zeros(4, dtype=torch.complex128, device=qf.backend.DEVICE)
    s = qf.backend.stream_handle()
    ptr = _native.ctypes.c_void_p(buf.data_ptr())
    assert lib.qfx_density_identity(s, ptr, -1, 1.0) == _native.QFX_ERANGE
    assert lib.qfx_density_trace(s, ptr, ptr, 16) == _native.QFX_ERANGE
    assert lib.qfx_density_outer(s, None, ptr, 1) == _native.QFX_EINVAL
    assert lib.qfx_density_partial_trace(s, ptr, ptr, 0, 1) == _native.QFX_ERANGE
    # a 0-qubit density is the 1 x 1 matrix of a scalar state
    assert lib.qfx_density_identity(s, ptr, 0, 0.5) == 0
    torch.cuda.synchronize()
    assert complex(buf[0].item()) == 0.5


def test_measure_reset_barrier_if_on_densities(qf):
    """Measure.evolve against the oracle with the same uniform draws;
    Reset.evolve raises TypeError as in the reference; Barrier passes the
    density through; If evolves on the recorded bit."""
    rng = np.random.default_rng(13)
    n = 4
    m0 = _rand_density(n, rng)
    for seed in range(6):
        np.random.seed(seed)
        rho = qf.Measure(2, 'c').evolve(qf.Density(m0.reshape([2] * (2 * n))))
        np.random.seed(seed)
        bit, ref = O.measure_density(m0, n, 2, np.random.random())
        assert rho.memory['c'] == bit
        assert np.max(np.abs(_mat(rho) - ref)) <= 1e-12
    with pytest.raises(TypeError):
        qf.Reset(0).evolve(rho)
    assert qf.Barrier(0, 1).evolve(rho) is rho
    flipped = qf.If(qf.X(0), 'c', rho.memory['c']).evolve(rho)
    ref = O.density_evolve(_mat(rho), O.gate_matrix('X'), n, [0])
    assert np.max(np.abs(_mat(flipped) - ref)) <= 1e-12


@pytest.mark.parametrize("traced", [[0], [5], [1, 4], [0, 2, 3], [5, 4, 3, 2, 1]])
def test_partial_trace_matches_oracle(qf, traced):
    rng = np.random.default_rng(20 + len(traced))
    n = 6
    m0 = _rand_density(n, rng)
    print("partial trace over", traced, flush=True)
    rho = qf.Density(m0.reshape([2] * (2 * n)), qubits=list('abcdef'))
    out = rho.partial_trace([rho.qubits[a] for a in traced])
    assert list(out.qubits) == [q for i, q in enumerate('abcdef') if i not in traced]
    ref = O.density_partial_trace(m0, n, traced)
    assert np.max(np.abs(_mat(out) - ref)) <= 1e-12
    with pytest.raises(ValueError):
        rho.partial_trace(list('abcdef'))


def test_channel_evolve_and_composition(qf):
    """Channel.evolve runs the superoperator as one gate on the row and
    column axes (dense 2-qubit, monomial 4-qubit and, for a 2-qubit Kraus
    channel, dense 4-qubit with the matrix in device memory); composition
    by @ equals evolving by each channel in turn."""
    rng = np.random.default_rng(17)
    n = 5
    m0 = _rand_density(n, rng)
    rho = qf.Density(m0.reshape([2] * (2 * n)))
    for gate, name, params, qubits in ((qf.RX(0.9, 2), 'RX', (0.9,), [2]),
                                       (qf.CNOT(4, 0), 'CNOT', (), [4, 0])):
        print("channel of", name, flush=True)
        out = gate.aschannel().evolve(rho)
        ref = O.density_evolve(m0, O.gate_matrix(name, *params), n, qubits)
        assert np.max(np.abs(_mat(out) - ref)) <= 1e-12
    chan = qf.Kraus([qf.RX(0.4, 3), qf.CNOT(3, 1), qf.H(1)], [0.5, 0.3, 0.2])
    out = chan.aschannel().evolve(rho)
    ops = [(O.gate_matrix('RX', 0.4), [3], 0.5), (O.gate_matrix('CNOT'), [3, 1], 0.3),
           (O.gate_matrix('H'), [1], 0.2)]
    assert np.max(np.abs(_mat(out) - O.kraus_evolve(m0, ops, n))) <= 1e-12
    a, b = qf.Depolarizing(0.1, 1).aschannel(), qf.Damping(0.2, 1).aschannel()
    both = (a @ b).evolve(rho)
    step = qf.Depolarizing(0.1, 1).evolve(qf.Damping(0.2, 1).evolve(rho))
    assert np.max(np.abs(_mat(both) - _mat(step))) <= 1e-12


def test_kraus_run_on_states(qf):
    """Kraus.run draws a branch with probability w_k |K_k psi|^2 and
    renormalises: full damping sends |1> to |0> on every draw, and a
    depolarizing channel with p = 0 leaves the state unchanged."""
    one = qf.State(np.array([0.0, 1.0]))
    for seed in range(3):
        np.random.seed(seed)
        out = qf.Damping(1.0, 0).run(one)
        assert np.allclose(out.tensor.cpu().numpy(), [1.0, 0.0], atol=1e-15)
    rng = np.random.default_rng(2)
    psi = rng.normal(size=[2] * 3) + 1j * rng.normal(size=[2] * 3)
    psi /= np.linalg.norm(psi)
    ket = qf.State(psi)
    out = qf.Depolarizing(0.0, 1).run(ket)
    assert np.max(np.abs(out.tensor.cpu().numpy() - psi)) <= 1e-15


def _oracle_circuit_evolve(m, n, circ_desc):
    for name, params, qubits in circ_desc:
        m = O.density_evolve(m, O.gate_matrix(name, *params), n, qubits)
    return m


_CIRC = [("H", (), [0]), ("CNOT", (), [0, 1]), ("RX", (0.3,), [2]), ("CCNOT", (), [2, 0, 1]),
         ("ISWAP", (), [1, 2]), ("T", (), [1]), ("CZ", (), [2, 0])]


def test_circuit_aschannel_and_dag_evolve(qf):
    """Circuit.aschannel (circuits.py:122-127) and DAGCircuit.evolve /
    aschannel (dagcircuit.py:86-96) on a 3-qubit density against the oracle's
    gate-by-gate evolution; the reference's CCNOT known answer
    (tests/test_dagcircuit.py:76-94)."""
    rng = np.random.default_rng(41)
    n = 3
    m0 = _rand_density(n, rng)
    ref = _oracle_circuit_evolve(m0, n, _CIRC)
    circ = qf.Circuit([getattr(qf, nm)(*p, *q) for nm, p, q in _CIRC])
    rho = qf.Density(m0.reshape([2] * (2 * n)))
    chan = circ.aschannel()
    assert chan.qubits == (0, 1, 2)
    assert np.max(np.abs(_mat(chan.evolve(rho)) - ref)) <= 1e-12
    dag = qf.DAGCircuit(circ)
    assert np.max(np.abs(_mat(dag.evolve(rho)) - ref)) <= 1e-12
    assert np.max(np.abs(_mat(dag.aschannel().evolve(rho)) - ref)) <= 1e-12
    # the superoperator of a unitary circuit is U (x) conj(U)
    u = O.circuit_unitary(_CIRC, (0, 1, 2))
    assert np.max(np.abs(chan.asoperator() - np.kron(u, u.conj()))) <= 1e-12
    # reference KAT: CCNOT against its decomposition, as a DAG
    rho1 = qf.CCNOT(0, 1, 2).evolve(rho)
    dag = qf.DAGCircuit(qf.ccnot_circuit([0, 1, 2]))
    assert np.max(np.abs(_mat(dag.evolve(rho)) - _mat(rho1))) <= 1e-12
    assert np.max(np.abs(_mat(dag.aschannel().evolve(rho)) - _mat(rho1))) <= 1e-12


def test_density_permute_stays_a_density(qf):
    rng = np.random.default_rng(42)
    n = 4
    m0 = _rand_density(n, rng)
    rho = qf.Density(m0.reshape([2] * (2 * n)), qubits=("a", "b", "c", "d"))
    order = ("c", "a", "d", "b")
    out = rho.permute(order)
    assert isinstance(out, qf.Density)
    assert out.qubits == order
    perm = [2, 0, 3, 1]
    want = m0.reshape([2] * (2 * n)).transpose(perm + [n + p for p in perm])
    assert np.array_equal(_mat(out), want.reshape(2 ** n, 2 ** n))
    assert abs(complex(out.trace().item()) - 1.0) <= 1e-12
    pt = out.partial_trace(["a"])
    ref = O.density_partial_trace(m0, n, [0])
    sub = ref.reshape([2] * 6).transpose([1, 2, 0, 4, 5, 3]).reshape(8, 8)   # (b,c,d) -> (c,d,b)
    assert np.max(np.abs(_mat(pt) - sub)) <= 1e-12

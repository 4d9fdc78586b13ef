"""The oracle's density-matrix functions, pinned by closed forms and by the
state-vector oracle (no GPU).

* Pure states: evolving |psi><psi| by U equals |U psi><U psi| with U psi
  from the state-vector oracle (tensormul, pinned in test_oracle.py).
* Kraus channels on one qubit have textbook closed forms: amplitude damping
  sends |1><1| to diag(p, 1 - p); depolarizing sends |0><0| to
  diag(1 - 2p/3, 2p/3); dephasing scales the coherences of |+><+| by 1 - p.
* Complete Kraus sets preserve trace and hermiticity; the maximally mixed
  state is a fixed point of every unitary.
"""
import numpy as np
import pytest

from oracle import qf_oracle as O


def _rand_unitary(k, rng):
    m = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
    q, r = np.linalg.qr(m)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _rand_density(n, rng):
    g = rng.normal(size=(2 ** n, 2 ** n)) + 1j * rng.normal(size=(2 ** n, 2 ** n))
    m = g @ g.conj().T
    return m / np.trace(m)


@pytest.mark.parametrize("k,idx", [(1, [2]), (2, [3, 0]), (3, [1, 4, 2])])
def test_pure_state_evolution_matches_state_vector_oracle(k, idx):
    rng = np.random.default_rng(k)
    n = 5
    psi = rng.normal(size=[2] * n) + 1j * rng.normal(size=[2] * n)
    psi /= np.linalg.norm(psi)
    U = _rand_unitary(k, rng)
    out = O.tensormul(U.reshape([2] * (2 * k)), psi, idx)
    ref = O.density_from_state(out)
    got = O.density_evolve(O.density_from_state(psi), U, n, idx)
    assert np.max(np.abs(got - ref)) < 1e-13


def test_damping_depolarizing_dephasing_closed_forms():
    p = 0.3
    one = np.diag([0.0, 1.0]).astype(complex)
    damp = [(np.array([[1, 0], [0, np.sqrt(1 - p)]]), [0], 1.0),
            (np.array([[0, np.sqrt(p)], [0, 0]]), [0], 1.0)]
    assert np.allclose(O.kraus_evolve(one, damp, 1), np.diag([p, 1 - p]), atol=1e-15)
    I2, X = np.eye(2), np.array([[0, 1], [1, 0]])
    Y, Z = np.array([[0, -1j], [1j, 0]]), np.diag([1.0, -1.0])
    zero = np.diag([1.0, 0.0]).astype(complex)
    dep = [(I2, [0], 1 - p), (X, [0], p / 3), (Y, [0], p / 3), (Z, [0], p / 3)]
    assert np.allclose(O.kraus_evolve(zero, dep, 1), np.diag([1 - 2 * p / 3, 2 * p / 3]),
                       atol=1e-15)
    plus = np.full((2, 2), 0.5, dtype=complex)
    deph = [(I2, [0], 1 - p / 2), (Z, [0], p / 2)]
    want = np.array([[0.5, 0.5 * (1 - p)], [0.5 * (1 - p), 0.5]])
    assert np.allclose(O.kraus_evolve(plus, deph, 1), want, atol=1e-15)


def test_trace_hermiticity_and_mixed_fixed_point():
    rng = np.random.default_rng(3)
    n = 4
    rho = _rand_density(n, rng)
    p = 0.2
    ops = [(np.array([[1, 0], [0, np.sqrt(1 - p)]]), [1], 1.0),
           (np.array([[0, np.sqrt(p)], [0, 0]]), [1], 1.0)]
    out = O.kraus_evolve(rho, ops, n)
    assert abs(np.trace(out) - 1.0) < 1e-13
    assert np.max(np.abs(out - out.conj().T)) < 1e-14
    mixed = np.eye(2 ** n) / 2 ** n
    U = _rand_unitary(2, rng)
    assert np.max(np.abs(O.density_evolve(mixed, U, n, [3, 1]) - mixed)) < 1e-15


def test_kraus_completeness_on_the_host():
    """kraus_iscomplete embeds K^H K with the oracle's full-operator
    convention and accepts exactly the complete weighted sets."""
    import quantumflow_amd as qf
    from quantumflow_amd.channels import _embed
    rng = np.random.default_rng(4)
    m = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    assert np.allclose(_embed(m, [2, 0], 3), O.full_operator(m, 3, [2, 0]))
    assert qf.kraus_iscomplete(qf.Damping(0.3, 0))
    assert qf.kraus_iscomplete(qf.Depolarizing(0.2, 1))
    assert qf.kraus_iscomplete(qf.Dephasing(0.1, 0))
    assert qf.kraus_iscomplete(qf.Kraus([qf.X(0), qf.CNOT(0, 1)], [0.5, 0.5]))
    assert not qf.kraus_iscomplete(qf.Kraus([qf.X(0), qf.CNOT(0, 1)], [0.5, 0.4]))


def test_superoperator_equals_kraus_sum():
    """Kraus.superoperator applied to vec(rho) (row-major) equals the
    oracle's sum of K rho K^H, for a channel whose operators act on
    different qubit subsets."""
    import quantumflow_amd as qf
    rng = np.random.default_rng(6)
    chan = qf.Kraus([qf.RX(0.4, 3), qf.CNOT(3, 1), qf.H(1)], [0.5, 0.3, 0.2])
    rho = _rand_density(2, rng)                 # qubits (1, 3) in that order
    got = (chan.superoperator() @ rho.reshape(-1)).reshape(4, 4)
    ops = [(O.gate_matrix('RX', 0.4), [1], 0.5), (O.gate_matrix('CNOT'), [1, 0], 0.3),
           (O.gate_matrix('H'), [0], 0.2)]
    assert np.max(np.abs(got - O.kraus_evolve(rho, ops, 2))) < 1e-14


def test_measure_density_closed_forms():
    """diag(p, 1-p): P0 rho P0 = diag(p, 0), whose squared Frobenius norm is
    p^2, so draws below p^2 give |0><0| and draws above give |1><1|; a
    coherent |+><+| loses its coherence either way."""
    p = 0.6
    rho = np.diag([p, 1 - p]).astype(complex)
    bit, out = O.measure_density(rho, 1, 0, p * p - 1e-9)
    assert bit == 0 and np.allclose(out, np.diag([1.0, 0.0]))
    bit, out = O.measure_density(rho, 1, 0, p * p + 1e-9)
    assert bit == 1 and np.allclose(out, np.diag([0.0, 1.0]))
    plus = np.full((4, 4), 0.25, dtype=complex)      # |++><++|
    bit, out = O.measure_density(plus, 2, 1, 0.01)
    assert bit == 0
    want = np.zeros((4, 4)); want[np.ix_([0, 2], [0, 2])] = 0.5
    assert np.allclose(out, want)


def test_partial_trace_closed_forms():
    """tr_B(rho_A (x) rho_B) = rho_A for any axes split; the Bell state
    traces to I/2; the trace of the reduced density is that of the whole."""
    rng = np.random.default_rng(8)
    ra, rb = _rand_density(1, rng), _rand_density(2, rng)
    rho = np.kron(ra, rb)                     # axes 0 | 1, 2
    assert np.allclose(O.density_partial_trace(rho, 3, [1, 2]), ra, atol=1e-14)
    assert np.allclose(O.density_partial_trace(rho, 3, [0]), rb, atol=1e-14)
    bell = np.zeros(4, dtype=complex)
    bell[0] = bell[3] = 1 / np.sqrt(2)
    b = O.density_from_state(bell)
    assert np.allclose(O.density_partial_trace(b, 2, [0]), np.eye(2) / 2, atol=1e-15)
    assert np.allclose(O.density_partial_trace(b, 2, [1]), np.eye(2) / 2, atol=1e-15)
    r4 = _rand_density(4, rng)
    assert abs(np.trace(O.density_partial_trace(r4, 4, [1, 3])) - 1.0) < 1e-13


def test_channel_superoperator_host_forms():
    """Gate.aschannel is U (x) conj(U); its Choi matrix is rank one with
    trace 2^N (a unitary channel); sharp is an involution; the channel of a
    Kraus set applied to vec(rho) equals the oracle's Kraus sum."""
    import quantumflow_amd as qf
    rng = np.random.default_rng(10)
    g = qf.CNOT(0, 1)
    ch = g.aschannel()
    U = O.gate_matrix('CNOT')
    assert np.allclose(ch.asoperator(), np.kron(U, U.conj()))
    choi = ch.choi()
    assert np.linalg.matrix_rank(choi, tol=1e-10) == 1
    assert abs(np.trace(choi) - 4.0) < 1e-12
    assert np.allclose(ch.sharp.sharp.asoperator(), ch.asoperator())
    damp = qf.Damping(0.3, 0).aschannel()
    rho = _rand_density(1, rng)
    ops = [(np.array([[1, 0], [0, np.sqrt(0.7)]]), [0], 1.0),
           (np.array([[0, np.sqrt(0.3)], [0, 0]]), [0], 1.0)]
    got = (damp.asoperator() @ rho.reshape(-1)).reshape(2, 2)
    assert np.allclose(got, O.kraus_evolve(rho, ops, 1))
    assert np.allclose((damp + damp * 0.5).asoperator(), 1.5 * damp.asoperator())


def test_join_channels_and_channel_to_kraus():
    """A joined channel acts on a product density as the two channels on
    their factors; converting a channel to Kraus operators and back gives
    the same superoperator, and the operators are complete."""
    import quantumflow_amd as qf
    rng = np.random.default_rng(12)
    a = qf.Damping(0.3, 0).aschannel()
    b = qf.CNOT(1, 2).aschannel()
    j = qf.join_channels(a, b)
    assert tuple(j.qubits) == (0, 1, 2)
    ra, rb = _rand_density(1, rng), _rand_density(2, rng)
    out = (j.asoperator() @ np.kron(ra, rb).reshape(-1)).reshape(8, 8)
    oa = (a.asoperator() @ ra.reshape(-1)).reshape(2, 2)
    ob = (b.asoperator() @ rb.reshape(-1)).reshape(4, 4)
    assert np.allclose(out, np.kron(oa, ob))
    for chan in (qf.Depolarizing(0.2, 0).aschannel(), qf.Damping(0.4, 1).aschannel(),
                 qf.Kraus([qf.RX(0.4, 1), qf.CNOT(1, 0)], [0.6, 0.4]).aschannel()):
        k = qf.channel_to_kraus(chan)
        assert np.allclose(k.aschannel().asoperator(), chan.asoperator())
        assert qf.kraus_iscomplete(k)

"""GPU sampling and mid-circuit measurement (qfx_sample, qfx_histogram,
qfx_qubit_probs; State.sample / measure / expectation(trials), Measure, Reset,
If, Barrier) against the oracle on the same inputs.

Device draws use a SplitMix64 stream keyed by a seed, not numpy's generator,
so counts match the reference's in distribution (chi-square) and the device
indices match the oracle's inverse-CDF restatement draw for draw, except for
draws whose uniform lands within rounding of a CDF boundary."""
import numpy as np
import pytest

from oracle import qf_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qf():
    import quantumflow_amd as qf
    assert qf.backend.gpu_available()
    return qf


def _random_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    v /= np.linalg.norm(v)
    return v


def _dev(qf, v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v.astype(np.complex128))).cuda()


def _check_draws(qf, psi, trials, seed):
    n = int(np.log2(psi.size))
    got = qf.backend.evaluate(qf.backend.sample_indices(_dev(qf, psi), trials, seed))
    want, margin = O.sample_icdf(psi.reshape([2] * n), trials, seed)
    total = np.sum(np.abs(psi) ** 2)
    bad = np.nonzero(got != want)[0]
    print(f"n={n} trials={trials} seed={seed}: {len(bad)} draws differ")
    assert np.all(margin[bad] < 1e-12 * total), (bad[:10], got[bad[:10]], want[bad[:10]])
    assert len(bad) <= max(2, trials // 10000)
    return got


@pytest.mark.parametrize("n", [1, 2, 5, 6, 7, 9, 10, 11, 13, 16, 20])
def test_draws_match_oracle(qf, n):
    psi = _random_state(n, 100 + n)
    _check_draws(qf, psi, 4096, 2024 + n)


def test_draws_sparse_and_zero_chunks(qf):
    # zero chunks (1024-amplitude blocks) between nonzero ones, zero lanes
    # inside a chunk, and an unnormalised state
    n = 14
    psi = _random_state(n, 7)
    psi[1024:3072] = 0
    psi[5000:5100] = 0
    psi[(np.arange(psi.size) % 17) == 3] = 0
    psi[-1024:] = 0
    got = _check_draws(qf, psi * 2.5, 20000, 77)
    assert np.all(np.abs(psi[got]) > 0)


def test_draws_basis_states_exact(qf):
    for n, k in [(1, 0), (1, 1), (6, 63), (10, 1023), (12, 1024), (18, 2 ** 18 - 1), (18, 77777)]:
        psi = np.zeros(2 ** n, complex)
        psi[k] = -1j
        got = qf.backend.evaluate(qf.backend.sample_indices(_dev(qf, psi), 1000, n * 31 + k))
        assert np.all(got == k), (n, k)


def test_draws_uniform_closed_form(qf):
    n = 12
    psi = np.full(2 ** n, 2 ** (-n / 2), complex)
    got = qf.backend.evaluate(qf.backend.sample_indices(_dev(qf, psi), 10000, 5))
    u = O.splitmix64_uniform(5, 10000)
    np.testing.assert_array_equal(got, np.floor(u * 2 ** n).astype(np.int64))


def test_draws_reproducible_and_seed_dependent(qf):
    psi = _dev(qf, _random_state(15, 3))
    a = qf.backend.evaluate(qf.backend.sample_indices(psi, 5000, 42))
    b = qf.backend.evaluate(qf.backend.sample_indices(psi, 5000, 42))
    c = qf.backend.evaluate(qf.backend.sample_indices(psi, 5000, 43))
    np.testing.assert_array_equal(a, b)
    assert np.mean(a != c) > 0.9


def test_empty_and_invalid(qf):
    psi = _dev(qf, _random_state(4, 1))
    assert qf.backend.sample_indices(psi, 0, 1).numel() == 0
    assert int(qf.backend.evaluate(qf.backend.sample_counts(psi, 0, 1)).sum()) == 0
    with pytest.raises(ValueError):
        qf.backend.sample_indices(psi, -1, 1)
    with pytest.raises(ValueError):
        qf.backend.qubit_probs(psi, 4)


def test_counts_histogram(qf):
    n = 11
    psi = _random_state(n, 9)
    d = _dev(qf, psi)
    idx = qf.backend.evaluate(qf.backend.sample_indices(d, 300000, 8))
    counts = qf.backend.evaluate(qf.backend.sample_counts(d, 300000, 8))
    assert counts.shape == (2,) * n
    np.testing.assert_array_equal(counts.ravel(), np.bincount(idx, minlength=2 ** n))
    p = np.abs(psi) ** 2
    mask = p * 300000 >= 5
    e = 300000 * p[mask]
    chi2 = np.sum((counts.ravel()[mask] - e) ** 2 / e)
    dof = mask.sum() - 1
    print(f"chi2 {chi2:.1f} dof {dof}")
    assert chi2 < dof + 6 * np.sqrt(2 * dof)


def test_large_state_sampled_z_matches_exact(qf):
    # size-independent property at 24 qubits: <Z_a> from 2^20 draws agrees
    # with the exact one-pass <Z_a> within 6 standard errors
    from quantumflow_amd.workloads import build_gates, random_mixed_circuit
    n = 24
    circ = qf.Circuit(build_gates(random_mixed_circuit(n, 10, seed=11)))
    ket = circ.run(qf.zero_state(n))
    trials = 1 << 20
    idx = qf.backend.evaluate(ket.sample_indices(trials, seed=2026))
    z_exact = qf.backend.evaluate(ket.expectation_z())
    for a in range(n):
        bits = (idx >> (n - 1 - a)) & 1
        z = np.mean(1 - 2 * bits)
        se = np.sqrt(max(1 - z_exact[a] ** 2, 1e-12) / trials)
        assert abs(z - z_exact[a]) < 6 * se + 1e-9, (a, z, z_exact[a])


def test_qubit_probs_vs_oracle(qf):
    for n in (1, 3, 12, 17):
        psi = _random_state(n, n) * 1.7
        d = _dev(qf, psi)
        bits_all = np.arange(2 ** n)
        for axis in range(n):
            p = qf.backend.evaluate(qf.backend.qubit_probs(d, axis))
            bits = (bits_all >> (n - 1 - axis)) & 1
            want0 = np.sum(np.abs(psi[bits == 0]) ** 2)
            want1 = np.sum(np.abs(psi[bits == 1]) ** 2)
            assert abs(p[0] - want0) < 1e-13 and abs(p[1] - want1) < 1e-13


# ---- the reference's own tests, on the device -------------------------------

def test_state_measure_kat(qf):
    # tests/test_states.py:108-121
    ket = qf.zero_state(2)
    assert np.allclose(ket.measure(), [0, 0])
    ket = qf.X(0).run(ket)
    assert np.allclose(ket.measure(), [1, 0])
    ket = qf.H(0).run(ket)
    ket = qf.CNOT(0, 1).run(ket)
    seen = set()
    for _ in range(40):
        res = ket.measure()
        assert res[0] == res[1]
        seen.add(int(res[0]))
    assert seen == {0, 1}


def test_state_sample_kat(qf):
    # tests/test_states.py:124-130
    ket = qf.zero_state(2)
    ket = qf.H(0).run(ket)
    ket = qf.CNOT(0, 1).run(ket)
    samples = ket.sample(10)
    assert samples.sum() == 10
    assert samples.shape == (2, 2)
    assert samples[0, 1] == 0 and samples[1, 0] == 0


def test_state_expectation_trials_kat(qf):
    # tests/test_states.py:133-141, plus the sampled value's spread
    ket = qf.zero_state(1)
    ket = qf.H(0).run(ket)
    m = ket.expectation([0.4, 0.6])
    assert abs(float(qf.backend.evaluate(m)) - 0.5) < 1e-12
    np.random.seed(1)
    m = float(qf.backend.evaluate(ket.expectation([0.4, 0.6], 10000)))
    assert abs(m - 0.5) < 6 * 0.1 / np.sqrt(10000)


def test_sample_seeded_by_numpy(qf):
    ket = qf.Circuit([qf.H(q) for q in range(6)]).run(qf.zero_state(6))
    np.random.seed(3)
    a = ket.sample(1000)
    np.random.seed(3)
    b = ket.sample(1000)
    np.testing.assert_array_equal(a, b)


def test_measure_op_vs_oracle(qf):
    n = 10
    psi = _random_state(n, 21)
    for axis in (0, 4, 9):
        for seed in range(4):
            ket = qf.State(psi.reshape([2] * n))
            np.random.seed(seed)
            out = qf.Measure(axis, ('c', axis)).run(ket)
            np.random.seed(seed)
            want, bit, _ = O.measure_qubit(psi.reshape([2] * n), axis, np.random.random())
            assert out.memory[('c', axis)] == bit
            err = np.max(np.abs(out.vec.asarray() - want))
            print(f"measure axis {axis} seed {seed}: bit {bit} err {err:.2e}")
            assert err < 1e-12


def test_measure_collapses_bell_pairs(qf):
    circ = qf.Circuit([qf.H(0), qf.CNOT(0, 1), qf.CNOT(1, 2),
                       qf.Measure(0, 'a'), qf.Measure(1, 'b'), qf.Measure(2, 'c')])
    outcomes = set()
    for seed in range(12):
        np.random.seed(seed)
        ket = circ.run()
        m = ket.memory
        assert m['a'] == m['b'] == m['c']
        outcomes.add(m['a'])
        z = qf.backend.evaluate(ket.expectation_z())
        np.testing.assert_allclose(z, [1 - 2 * m['a']] * 3, atol=1e-12)
    assert outcomes == {0, 1}


def test_reset_vs_oracle(qf):
    n = 9
    psi = _random_state(n, 31)
    for qubits in ((), (0,), (3, 8), (1, 2, 5)):
        ket = qf.State(psi.reshape([2] * n))
        out = qf.Reset(*qubits).run(ket)
        axes = list(qubits) if qubits else list(range(n))
        want = O.reset_qubits(psi.reshape([2] * n), axes)
        err = np.max(np.abs(out.vec.asarray() - want))
        assert err < 1e-12, (qubits, err)


def test_if_and_barrier_kat(qf):
    # tests/test_stdops.py:24-63 (state-vector runs)
    circ = qf.Circuit()
    circ += qf.Barrier(0, 1, 2)
    circ += qf.Barrier(0, 1, 2).H
    circ.run()
    assert str(qf.Barrier(0, 1, 2)) == 'BARRIER 0 1 2'
    c = [('c', 0), ('c', 1)]
    for m1, value, want in ((1, True, 1), (0, True, 0), (0, False, 1)):
        circ = qf.Circuit([qf.Move(c[0], 0), qf.Move(c[1], m1),
                           qf.If(qf.X(0), c[1], value=value), qf.Measure(0, c[0])])
        ket = circ.run()
        assert ket.memory[c[0]] == want


def test_fused_segments_around_measure_and_reset(qf):
    # a 16-qubit circuit takes the fused executor for its gate runs; the
    # measurement and reset between them must see the same state as the
    # per-gate path
    from quantumflow_amd.workloads import build_gates, random_mixed_circuit
    n = 16
    g = build_gates(random_mixed_circuit(n, 12, seed=5))
    h = len(g) // 2
    circ = qf.Circuit(g[:h] + [qf.Measure(3, 'm'), qf.Reset(5)] + g[h:] + [qf.Measure(9, 'k')])
    for seed in range(3):
        np.random.seed(seed)
        a = circ.run(qf.zero_state(n), fused=True)
        np.random.seed(seed)
        b = circ.run(qf.zero_state(n), fused=False)
        assert a.memory == b.memory
        err = np.max(np.abs(a.vec.asarray() - b.vec.asarray()))
        print(f"seed {seed}: memory {dict(a.memory)} max |d amp| {err:.2e}")
        assert err < 1e-12

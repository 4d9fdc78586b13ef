"""Pins of the oracle's sampling and measurement functions
(oracle/qf_oracle.py: splitmix64, sample_icdf, sample_counts_reference,
measure_reference, measure_qubit, reset_qubits) by published vectors, closed
forms and brute force.  CPU only."""
import numpy as np
import pytest

from oracle import qf_oracle as O


def _random_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    v /= np.linalg.norm(v)
    return v.reshape([2] * n)


def test_splitmix64_published_vectors():
    # First outputs of SplitMix64 started at state 0, as listed by the
    # generator's published reference implementations.
    assert [O.splitmix64(0, t) for t in range(4)] == [
        16294208416658607535, 7960286522194355700,
        487617019471545679, 17909611376780542444]


def test_uniform_range_and_moments():
    u = O.splitmix64_uniform(12345, 20000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / 20000)
    assert abs(u.var() - 1 / 12) < 0.005


def test_icdf_basis_state_is_exact():
    for n, k in [(1, 1), (3, 5), (6, 0), (6, 63), (11, 1234)]:
        psi = np.zeros(2 ** n, complex)
        psi[k] = 1j
        idx, _ = O.sample_icdf(psi.reshape([2] * n), 500, 7)
        assert np.all(idx == k)


def test_icdf_uniform_state_closed_form():
    # |+>^n: cdf_i = (i+1)/2^n exactly, so draw t is floor(u_t 2^n)
    for n in (1, 4, 9):
        psi = np.full(2 ** n, 2 ** (-n / 2), complex)
        idx, _ = O.sample_icdf(psi.reshape([2] * n), 3000, 99)
        u = O.splitmix64_uniform(99, 3000)
        np.testing.assert_array_equal(idx, np.floor(u * 2 ** n).astype(np.int64))


def test_icdf_two_point():
    # probabilities {1/4 at index 2, 3/4 at index 5} of 8: u < 1/4 -> 2
    psi = np.zeros(8, complex)
    psi[2] = 0.5
    psi[5] = np.sqrt(0.75)
    idx, _ = O.sample_icdf(psi.reshape([2] * 3), 4000, 3)
    u = O.splitmix64_uniform(3, 4000)
    np.testing.assert_array_equal(idx, np.where(u < 0.25, 2, 5))


def test_icdf_never_draws_zero_probability():
    psi = _random_state(8, 1).ravel()
    psi[::3] = 0
    idx, _ = O.sample_icdf(psi.reshape([2] * 8), 20000, 5)
    assert np.all(np.abs(psi[idx]) > 0)


def test_icdf_unnormalised_state_same_distribution():
    psi = _random_state(5, 2)
    a, _ = O.sample_icdf(psi, 1000, 11)
    b, _ = O.sample_icdf(psi * 3.0, 1000, 11)
    np.testing.assert_array_equal(a, b)


def _chi2_ok(counts, probs, trials):
    mask = probs * trials >= 5
    chi2 = np.sum((counts[mask] - trials * probs[mask]) ** 2 / (trials * probs[mask]))
    dof = mask.sum() - 1
    return chi2 < dof + 6 * np.sqrt(2 * dof)


def test_icdf_and_reference_sampler_agree_in_distribution():
    psi = _random_state(6, 3)
    p = O.probabilities(psi).ravel()
    trials = 200000
    idx, _ = O.sample_icdf(psi, trials, 2024)
    assert _chi2_ok(np.bincount(idx, minlength=64), p, trials)
    np.random.seed(0)
    counts = O.sample_counts_reference(psi, trials).ravel()
    assert counts.sum() == trials
    assert _chi2_ok(counts, p, trials)


def test_measure_reference_kat():
    # tests/test_states.py:108-121
    ket = O.zero_state(2)
    assert list(O.measure_reference(ket)) == [0, 0]
    x = O.gate_matrix('X')
    ket = O.tensormul(x.reshape([2] * 2), ket, [0])
    assert list(O.measure_reference(ket)) == [1, 0]
    bell = np.zeros(4, complex)
    bell[0] = bell[3] = np.sqrt(0.5)
    for _ in range(20):
        r = O.measure_reference(bell.reshape(2, 2))
        assert r[0] == r[1]


def test_measure_qubit_bell_closed_form():
    bell = np.zeros(4, complex)
    bell[0] = bell[3] = np.sqrt(0.5)
    bell = bell.reshape(2, 2)
    for axis in (0, 1):
        post, bit, p0 = O.measure_qubit(bell, axis, 0.3)
        assert bit == 0 and abs(p0 - 0.5) < 1e-15
        np.testing.assert_allclose(post.ravel(), [1, 0, 0, 0], atol=1e-15)
        post, bit, _ = O.measure_qubit(bell, axis, 0.7)
        assert bit == 1
        np.testing.assert_allclose(post.ravel(), [0, 0, 0, 1], atol=1e-15)


def test_measure_qubit_brute_force():
    n = 5
    psi = _random_state(n, 4)
    flat = psi.ravel()
    for axis in range(n):
        bits = (np.arange(2 ** n) >> (n - 1 - axis)) & 1
        p0 = np.sum(np.abs(flat[bits == 0]) ** 2)
        for rand, want in ((p0 * 0.999, 0), (p0 * 1.001 + 1e-9, 1)):
            post, bit, prob_zero = O.measure_qubit(psi, axis, rand)
            assert bit == want
            assert abs(prob_zero - p0) < 1e-14
            expect = np.where(bits == want, flat, 0)
            expect = expect / np.linalg.norm(expect)
            np.testing.assert_allclose(post.ravel(), expect, atol=1e-14)


def test_reset_closed_forms():
    n = 4
    psi = _random_state(n, 5)
    # all qubits: [[1,1],[0,0]]^{(x) n} psi = (sum_z psi_z) |0..0>
    out = O.reset_qubits(psi, range(n)).ravel()
    s = psi.sum()
    assert abs(out[0] - s / abs(s)) < 1e-14
    assert np.allclose(out[1:], 0)
    # one qubit of a product state |a> (x) |b>: |0> (x) |b>, up to the
    # phase of a0 + a1
    a = np.array([0.6, 0.8j])
    b = _random_state(3, 6).ravel()
    prod = np.kron(a, b).reshape([2] * 4)
    out = O.reset_qubits(prod, [0]).ravel()
    ph = (a[0] + a[1]) / abs(a[0] + a[1])
    np.testing.assert_allclose(out, np.kron([1, 0], b) * ph, atol=1e-14)


def test_reset_planted_mistake_detected():
    # a reset that forgets to normalise fails the closed form above
    psi = _random_state(3, 7)
    r = np.array([[1, 1], [0, 0]], complex)
    t = psi
    for a in range(3):
        t = O.tensormul(r, t, [a])
    assert abs(abs(t.ravel()[0]) - 1) > 1e-3

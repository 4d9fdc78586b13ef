"""GPU parity of the per-gate kernels and readouts against the oracle.

Tolerances: dense and diagonal gates max|d amplitude| <= 1e-12 (complex128);
permutation gates (X, CNOT, SWAP, CCNOT, CSWAP) bitwise equal; cut tables
bitwise equal.
"""
import numpy as np
import pytest

from oracle import qf_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def bk():
    import quantumflow_amd.backend as bk
    assert bk.gpu_available(), "GPU visible but libqfx.so not loaded"
    return bk


def rand_state(n, rng):
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    return (v / np.linalg.norm(v)).reshape([2] * n)


def rand_unitary(k, rng):
    a = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
    q, r = np.linalg.qr(a)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def gpu_tensormul(bk, mat, psi, axes):
    k = len(axes)
    out = bk.tensormul(mat.reshape([2] * (2 * k)), bk.astensorproduct(psi), axes)
    return bk.evaluate(out)


@pytest.mark.parametrize("n", [5, 11, 14])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_dense_all_positions(bk, n, k):
    print("dense n=%d k=%d" % (n, k), flush=True)
    rng = np.random.default_rng(100 * n + k)
    psi = rand_state(n, rng)
    # every axis appears as a target, low and high bits included
    for start in range(0, n, 1 if n <= 5 else 3):
        axes = [(start + j * (1 + start % 2)) % n for j in range(k)]
        if len(set(axes)) < k:
            continue
        u = rand_unitary(k, rng)
        ref = O.tensormul(u.reshape([2] * (2 * k)), psi, axes)
        got = gpu_tensormul(bk, u, psi, axes)
        assert np.max(np.abs(got - ref)) <= TOL


@pytest.mark.parametrize("k", [4, 5])
def test_dense_device_matrix(bk, k):
    rng = np.random.default_rng(7 + k)
    n = 12
    psi = rand_state(n, rng)
    axes = list(rng.permutation(n)[:k])
    u = rand_unitary(k, rng)
    ref = O.tensormul(u.reshape([2] * (2 * k)), psi, axes)
    got = gpu_tensormul(bk, u, psi, axes)
    assert np.max(np.abs(got - ref)) <= TOL


@pytest.mark.parametrize("name,params,k", [
    ('RZ', (0.731,), 1), ('PHASE', (2.1,), 1), ('CZ', (), 2), ('ZZ', (0.37,), 2),
    ('CPHASE', (1.3,), 2), ('S', (), 1), ('T', (), 1)])
def test_diagonal_gates(bk, name, params, k):
    rng = np.random.default_rng(11)
    n = 13
    psi = rand_state(n, rng)
    m = O.gate_matrix(name, *params)
    for axes in ([0, 1][:k], [n - 1, 0][:k], [5, 9][:k], [n - 2, n - 1][:k]):
        ref = O.tensormul(m.reshape([2] * (2 * k)), psi, axes)
        got = gpu_tensormul(bk, m, psi, axes)
        assert np.max(np.abs(got - ref)) <= TOL


@pytest.mark.parametrize("name,k", [('X', 1), ('CNOT', 2), ('SWAP', 2), ('CCNOT', 3),
                                    ('CSWAP', 3), ('ISWAP', 2), ('Y', 1)])
def test_permutation_gates_exact(bk, name, k):
    rng = np.random.default_rng(12)
    n = 12
    psi = rand_state(n, rng)
    m = O.gate_matrix(name)
    for trial in range(6):
        axes = [int(a) for a in rng.permutation(n)[:k]]
        ref = O.tensormul(m.reshape([2] * (2 * k)), psi, axes)
        got = gpu_tensormul(bk, m, psi, axes)
        assert np.array_equal(got, ref), (name, axes)


def test_inplace_sparse_matches_out_of_place(bk):
    import torch
    rng = np.random.default_rng(13)
    n = 12
    psi = rand_state(n, rng)
    for name, params, k in [('CNOT', (), 2), ('CZ', (), 2), ('CCNOT', (), 3), ('PHASE', (0.4,), 1),
                            ('SWAP', (), 2), ('ISWAP', (), 2), ('RZ', (1.1,), 1)]:
        m = O.gate_matrix(name, *params)
        kind = bk.classify(m)
        axes = [int(a) for a in rng.permutation(n)[:k]]
        bits = [n - 1 - a for a in axes]
        src = bk.as_device(bk.astensorproduct(psi))
        out = torch.empty_like(src)
        bk.apply_gate(out, src, n, bits, kind)
        work = src.clone()
        bk.apply_gate(work, work, n, bits, kind, inplace_sparse=True)
        assert torch.equal(out, work), name


def test_readouts(bk):
    rng = np.random.default_rng(14)
    for n in (1, 3, 9, 17):
        a = rand_state(n, rng)
        b = rand_state(n, rng)
        ta, tb = bk.astensorproduct(a), bk.astensorproduct(b)
        assert abs(bk.evaluate(bk.inner(ta, tb)) - np.vdot(a, b)) <= TOL
        assert np.max(np.abs(bk.evaluate(bk.absolute(ta)) - np.absolute(a))) <= 1e-15
        assert np.max(np.abs(bk.evaluate(bk.probabilities(ta)) - O.probabilities(a))) <= 1e-15
        d = rng.uniform(-3, 3, size=[2] * n)
        assert abs(bk.evaluate(bk.expectation(d, ta)) - O.expectation(a, d)) <= TOL
        assert np.max(np.abs(bk.evaluate(bk.expectation_z(ta)) - O.expectation_z(a))) <= TOL
        assert abs(bk.evaluate(bk.norm(2 * ta)) - O.norm(2 * a)) <= 4 * TOL


def test_reductions_deterministic(bk):
    rng = np.random.default_rng(15)
    a = bk.astensorproduct(rand_state(20, rng))
    r1 = bk.evaluate(bk.inner(a, a))
    r2 = bk.evaluate(bk.inner(a, a))
    assert r1 == r2
    z1 = bk.evaluate(bk.expectation_z(a))
    z2 = bk.evaluate(bk.expectation_z(a))
    assert np.array_equal(z1, z2)


def test_permute_qubits(bk):
    from quantumflow_amd.qubits import permute_axes
    rng = np.random.default_rng(16)
    n = 9
    a = rand_state(n, rng)
    perm = [int(x) for x in rng.permutation(n)]
    got = bk.evaluate(permute_axes(bk.astensorproduct(a), perm))
    assert np.array_equal(got, np.transpose(a, perm))


def test_cut_table_bitwise(bk):
    import networkx as nx
    from quantumflow_amd.qaoa import graph_cuts
    rng = np.random.default_rng(17)
    for n in (3, 8, 14):
        g = nx.gnp_random_graph(n, 0.5, seed=int(rng.integers(1000)))
        edges = list(g.edges())
        assert np.array_equal(bk.evaluate(graph_cuts(g)), O.graph_cuts(n, edges))
        for (u, v) in edges:
            g[u][v]['weight'] = float(rng.uniform(0.1, 2))
        w = [g[u][v]['weight'] for u, v in edges]
        assert np.array_equal(bk.evaluate(graph_cuts(g)), O.graph_cuts(n, edges, w))


def test_errors_map_to_reference_types(bk):
    import quantumflow_amd as qf
    ket = qf.zero_state(3)
    with pytest.raises(ValueError):
        qf.X(7).run(ket)                       # unknown qubit (ops.py:164)
    with pytest.raises(AssertionError):
        bk.tensormul(np.eye(4).reshape([2] * 4), ket.tensor, [0])   # K mismatch
    with pytest.raises(ValueError):
        bk.tensormul(np.eye(4).reshape([2] * 4), ket.tensor, [1, 1])  # repeated target


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("T,L,block,width,mode", [
    (6, 3, 64, 1, 'lds'), (8, 3, 256, 3, 'lds'), (10, 4, 256, 4, 'lds'),
    (11, 3, 512, 4, 'lds'), (12, 5, 1024, 3, 'lds'), (13, 3, 256, 5, 'lds'),
    (13, 4, 512, 4, 'lds'), (6, 3, 64, 1, 'reg'), (8, 2, 256, 1, 'reg'),
    (11, 3, 256, 1, 'reg'), (12, 3, 256, 1, 'reg'), (12, 4, 512, 1, 'reg'),
    (13, 3, 512, 1, 'reg'), (9, 3, 256, 3, 'mfma'), (10, 3, 64, 4, 'mfma'),
    (12, 3, 256, 5, 'mfma'), (12, 3, 512, 4, 'mfma'), (13, 3, 256, 3, 'mfma'),
    (13, 4, 512, 5, 'mfma')])
def test_tile_plan_configs(bk, T, L, block, width, mode, persistent):
    """Every tiled-executor configuration reproduces the oracle's state."""
    import quantumflow_amd as qf
    from quantumflow_amd.fusion import TilePlan
    from quantumflow_amd.workloads import random_mixed_circuit, build_gates
    n = 15
    print("tile T=%d L=%d block=%d width=%d mode=%s persistent=%s"
          % (T, L, block, width, mode, persistent), flush=True)
    desc = random_mixed_circuit(n, 10, seed=T * 7 + L)
    ref = O.run_circuit(desc, qubits=tuple(range(n)))
    plan = TilePlan(build_gates(desc), tuple(range(n)), tile_bits=T, low_bits=L,
                    block=block, fuse_width=width, mode=mode, persistent=persistent)
    got = plan.run(qf.zero_state(n)).vec.asarray()
    assert np.max(np.abs(got - ref)) <= TOL


def test_tile_plan_permutations_exact(bk):
    """Tiled passes of permutation gates move amplitudes bit for bit."""
    import quantumflow_amd as qf
    from quantumflow_amd.fusion import TilePlan
    rng = np.random.default_rng(21)
    n = 14
    desc = []
    for _ in range(60):
        k = int(rng.integers(1, 4))
        name = {1: 'X', 2: ('CNOT', 'SWAP')[int(rng.integers(2))], 3: ('CCNOT', 'CSWAP')[int(rng.integers(2))]}[k]
        desc.append((name, (), tuple(int(q) for q in rng.permutation(n)[:k])))
    psi = rand_state(n, rng)
    ref = O.run_circuit(desc, qubits=tuple(range(n)), ket=psi)
    from quantumflow_amd.workloads import build_gates
    for mode, width in (('lds', 1), ('reg', 1), ('mfma', 4)):
        plan = TilePlan(build_gates(desc), tuple(range(n)), tile_bits=10, low_bits=3,
                        fuse_width=width, mode=mode)
        got = plan.run(qf.State(psi)).vec.asarray()
        # permutation blocks on the matrix cores multiply by exact 0/1
        assert np.array_equal(got, ref), mode


@pytest.mark.parametrize("flips", [True, False])
def test_lane_flips_on_device(bk, flips):
    """Plans whose selector-controlled X gates are lane flips (no amplitudes
    moved; folded tables, flagged reads, flipped stores), and the same plans without
    flips, reproduce the oracle's state on random circuits of 15 and 16
    qubits, for every register-phase kernel the plans use."""
    import quantumflow_amd as qf
    from quantumflow_amd import fusion
    from quantumflow_amd.fusion import TilePlan
    from quantumflow_amd.workloads import random_mixed_circuit, build_gates
    rng = np.random.default_rng(33)
    for n, T, block in ((15, 10, 64), (16, 12, 256)):
        for seed in range(3):
            print("lane flips=%s n=%d T=%d seed=%d" % (flips, n, T, seed), flush=True)
            desc = random_mixed_circuit(n, 16, seed=100 + seed)
            psi = rand_state(n, rng)
            ref = O.run_circuit(desc, qubits=tuple(range(n)), ket=psi)
            plan = TilePlan(build_gates(desc), tuple(range(n)), tile_bits=T, low_bits=3,
                            block=block, fuse_width=1, mode='reg', lane_flips=flips)
            nflip = sum(bool(plan.srcs[o.src + 4 * q] >> 8)
                        for o in plan.ops_host if o.kind == fusion.OP_RPHASE
                        for q in range(o.coef))
            assert (nflip > 0) == flips
            got = plan.run(qf.State(psi)).vec.asarray()
            assert np.max(np.abs(got - ref)) <= TOL


def test_tile_pass_of_only_absorbed_swaps(bk):
    """A pass left with no ops (its SWAP gates absorbed as a store
    permutation) still moves the amplitudes, bit for bit."""
    import quantumflow_amd as qf
    from quantumflow_amd.fusion import TilePlan
    from quantumflow_amd.workloads import build_gates
    n = 14
    desc = [('SWAP', (), (0, 1)), ('SWAP', (), (1, 3))]
    rng = np.random.default_rng(4)
    psi = rand_state(n, rng)
    ref = O.run_circuit(desc, qubits=tuple(range(n)), ket=psi)
    for persistent in (False, True):
        plan = TilePlan(build_gates(desc), tuple(range(n)), tile_bits=12, low_bits=3,
                        block=256, fuse_width=1, mode='reg', persistent=persistent)
        assert any(p[1][2] == 0 for p in plan.passes)
        got = plan.run(qf.State(psi)).vec.asarray()
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("kind", ["diagonal", "monomial"])
def test_sparse_4q_gates_two_groups_per_lane(bk, kind):
    """4- and 5-qubit diagonal and monomial gates on states large enough
    that every lane of the sparse kernel owns two amplitude groups (n = 13
    and 15: 2^(n-k) groups exceed the block times two), in place and out of
    place, against the oracle."""
    import torch
    rng = np.random.default_rng(22)
    for n in (13, 15):
        psi = rand_state(n, rng)
        for k in (4, 5):
            d = 2 ** k
            if kind == "diagonal":
                m = np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, size=d)))
            else:
                perm = rng.permutation(d)
                m = np.zeros((d, d), dtype=complex)
                m[np.arange(d), perm] = np.exp(1j * rng.uniform(0, 2 * np.pi, size=d))
            assert bk.classify(m).kind == (bk.GateKind.DIAGONAL if kind == "diagonal"
                                           else bk.GateKind.MONOMIAL)
            for trial in range(3):
                axes = [int(a) for a in rng.permutation(n)[:k]]
                ref = O.tensormul(m.reshape([2] * (2 * k)), psi, axes)
                got = gpu_tensormul(bk, m, psi, axes)
                assert np.max(np.abs(got - ref)) <= TOL, (n, k, axes)
                bits = [n - 1 - a for a in axes]
                work = bk.as_device(bk.astensorproduct(psi)).clone()
                bk.apply_gate(work, work, n, bits, bk.classify(m), inplace_sparse=True)
                got_ip = work.cpu().numpy().reshape([2] * n)
                assert np.max(np.abs(got_ip - ref)) <= TOL, (n, k, axes, "in place")
                torch.cuda.synchronize()


def test_density_cnot_7q(bk):
    """A 7-qubit density (a 14-axis tensor, 2^10 four-qubit groups of the
    U (x) conj U monomial pass, two per lane) through CNOT gates, against
    the oracle."""
    import quantumflow_amd as qf
    rng = np.random.default_rng(23)
    n = 7
    g = rng.normal(size=(2 ** n, 2 ** n)) + 1j * rng.normal(size=(2 ** n, 2 ** n))
    m = g @ g.conj().T
    m /= np.trace(m)
    rho = qf.Density(m.reshape([2] * (2 * n)))
    for (c, t) in ((0, 6), (5, 2), (3, 4)):
        m = O.density_evolve(m, O.gate_matrix('CNOT'), n, [c, t])
        rho = qf.CNOT(c, t).evolve(rho)
        got = rho.asoperator().cpu().numpy().reshape(2 ** n, 2 ** n)
        assert np.array_equal(got, m), (c, t)

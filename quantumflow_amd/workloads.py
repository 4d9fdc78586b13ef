"""Seeded synthetic workloads: the random circuits and graph ensembles the
benchmarks and parity tests run.  Circuits are returned as descriptions
[(gate name, params, qubits)] so the same circuit can be built as
quantumflow_amd Gates (``build_gates``) or handed to any other simulator.
"""

import networkx as nx
import numpy as np

ONEQ = ('RX', 'RY', 'RZ')


def layered_rotation_circuit(n=12, layers=20, seed=0):
    """Per layer: RX, RY or RZ (uniform) with angle U[0, 2pi) on every qubit,
    then the qubits shuffled into n/2 disjoint pairs, each given CNOT or CZ.
    12 qubits x 20 layers = 360 gates."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(layers):
        for q in range(n):
            name = ONEQ[rng.integers(3)]
            out.append((name, (float(rng.uniform(0, 2 * np.pi)),), (q,)))
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            name = ('CNOT', 'CZ')[rng.integers(2)]
            out.append((name, (), (int(perm[i]), int(perm[i + 1]))))
    return out


def random_mixed_circuit(n=28, depth=40, seed=0):
    """Per layer: shuffle the qubits, then repeatedly draw a gate class
    (1q 50%: RX/RY/RZ/PHASE with angle U[0, 2pi); 2q 45%: CNOT/CZ/ISWAP/SWAP;
    3q 5%: CCNOT) consuming qubits from the shuffled list, falling back to a
    1q gate when too few remain.  About 18 gates per layer at n=28."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(depth):
        order = [int(q) for q in rng.permutation(n)]
        pos = 0
        while pos < n:
            u = rng.random()
            k = 1 if u < 0.50 else (2 if u < 0.95 else 3)
            if pos + k > n:
                k = 1
            qs = tuple(order[pos:pos + k])
            pos += k
            if k == 1:
                name = ('RX', 'RY', 'RZ', 'PHASE')[rng.integers(4)]
                out.append((name, (float(rng.uniform(0, 2 * np.pi)),), qs))
            elif k == 2:
                out.append((('CNOT', 'CZ', 'ISWAP', 'SWAP')[rng.integers(4)], (), qs))
            else:
                out.append(('CCNOT', (), qs))
    return out


def build_gates(desc):
    """Build quantumflow_amd Gates from a circuit description."""
    from . import stdgates
    gates = []
    for name, params, qubits in desc:
        gates.append(stdgates.STDGATES[name](*params, *qubits))
    return gates


def er_graph(n, p_edge, seed):
    """G(n, p) with nodes 0..n-1 (networkx gnp_random_graph)."""
    return nx.gnp_random_graph(n, p_edge, seed=int(seed))


def qaoa_batch_instances(count=512, n_lo=10, n_hi=22, p=4, seed=0):
    """ER graphs with n ~ U{n_lo..n_hi}, edge probability ~ U[0.3, 0.6],
    unit weights; angles gamma ~ U[0, pi], beta ~ U[0, pi/2] per layer."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        pe = float(rng.uniform(0.3, 0.6))
        g = er_graph(n, pe, int(rng.integers(1 << 30)))
        gamma = rng.uniform(0, np.pi, size=p)
        beta = rng.uniform(0, np.pi / 2, size=p)
        out.append((g, beta, gamma))
    return out


def regular_instances(count=64, n_lo=16, n_hi=24, p=8, seed=0, degree=3):
    """Random d-regular graphs with even n ~ U{n_lo..n_hi} and angles as in
    ``qaoa_batch_instances``."""
    rng = np.random.default_rng(seed)
    evens = [n for n in range(n_lo, n_hi + 1) if (n * degree) % 2 == 0]
    out = []
    for i in range(count):
        n = int(evens[rng.integers(len(evens))])
        g = nx.random_regular_graph(degree, n, seed=int(rng.integers(1 << 30)))
        gamma = rng.uniform(0, np.pi, size=p)
        beta = rng.uniform(0, np.pi / 2, size=p)
        out.append((g, beta, gamma))
    return out


def gops_circuit(n=16, gates=100, seed=0):
    """The gate-operations-per-second workload (tools/benchmark.py:31-46):
    H on every qubit, then (gates - n) // 3 triples X(q0), T(q1),
    CNOT(q0, q1) on a random pair of distinct qubits.  The reference draws
    the pairs with Python's random.sample; here a seeded numpy generator."""
    rng = np.random.default_rng(seed)
    out = [('H', (), (q,)) for q in range(n)]
    for _ in range((gates - n) // 3):
        q0, q1 = (int(q) for q in rng.choice(n, size=2, replace=False))
        out += [('X', (), (q0,)), ('T', (), (q1,)), ('CNOT', (), (q0, q1))]
    return out

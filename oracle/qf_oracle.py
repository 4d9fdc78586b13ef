"""numpy restatement of QuantumFlow's state-vector hot path (TEST ORACLE).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  Never shipped or measured as the product.

Conventions restated from the reference:
  * a state of N qubits is a complex128 array of shape [2]*N; axis i belongs
    to qubit label ``qubits[i]``; the flat index is big-endian in axis order
    (quantumflow/qubits.py:63-68, tests/test_gates.py:59-77);
  * a K-qubit gate is a [2]*(2K) tensor ordered [ket_out..., bra_in...], i.e.
    a 2^K x 2^K matrix with row = output and gate qubit 0 as the MSB
    (quantumflow/qubits.py:70-80).

Each function cites the reference lines it follows.  How each one is pinned
(known-answer vectors, closed forms, brute force) is listed in
tests/test_oracle.py.
"""

import itertools
import math
import string
from math import pi, sqrt

import numpy as np

CTYPE = np.complex128
FTYPE = np.float64

# quantumflow/backend/numpybk.py:72-76
EINSUM_SUBSCRIPTS = string.ascii_lowercase + string.ascii_uppercase


# --------------------------------------------------------------------------
# Backend arithmetic (quantumflow/backend/numpybk.py)
# --------------------------------------------------------------------------

def astensor(array):
    """numpybk.py:97-102: coerce to complex128."""
    return np.asarray(array, dtype=CTYPE)


def astensorproduct(array):
    """numpybk.py:105-112: coerce and reshape to [2]*N."""
    tensor = astensor(array)
    n = int(math.log2(tensor.size))
    return tensor.reshape([2] * n)


def tensormul(tensor0, tensor1, indices):
    """numpybk.py:159-214: contract gate ``tensor0`` (rank 2K) into
    ``tensor1`` (rank N) on axes ``indices`` with one einsum call.

    Subscripts: output = first N letters; the gate's bra letters are the next
    K letters; the state's subscripts are the output letters with the
    ``indices`` positions replaced by the bra letters (numpybk.py:199-210).
    """
    n = tensor1.ndim
    k = tensor0.ndim // 2
    assert k == len(indices)
    out = list(EINSUM_SUBSCRIPTS[0:n])
    left_in = list(EINSUM_SUBSCRIPTS[n:n + k])
    left_out = [out[idx] for idx in indices]
    right = list(EINSUM_SUBSCRIPTS[0:n])
    for idx, s in zip(indices, left_in):
        right[idx] = s
    subscripts = ''.join(left_out + left_in + [','] + right + ['->'] + out)
    return np.einsum(subscripts, tensor0, tensor1)


def tensormul_matmul(tensor0, tensor1, indices):
    """Independent restatement of tensormul by transpose -> matmul ->
    inverse transpose, the scheme of the reference's TensorFlow/torch
    backends (quantumflow/backend/tensorflowbk.py:134-152,
    quantumflow/backend/torchbk.py:243-261).  Used only to cross-check
    ``tensormul``."""
    n = tensor1.ndim
    k = tensor0.ndim // 2
    assert k == len(indices)
    rest = [i for i in range(n) if i not in indices]
    perm = list(indices) + rest
    t1 = np.transpose(tensor1, perm).reshape(2 ** k, -1)
    g = tensor0.reshape(2 ** k, 2 ** k)
    res = (g @ t1).reshape([2] * n)
    inv = np.argsort(perm)
    return np.transpose(res, inv)


def full_operator(matrix, n, indices):
    """Brute force: the full 2^n x 2^n operator of a K-qubit gate acting on
    axes ``indices`` of an n-qubit state, built entry by entry from the
    index convention (quantumflow/qubits.py:63-80).  Only for n <= 10."""
    assert n <= 10
    k = len(indices)
    matrix = np.asarray(matrix, dtype=CTYPE).reshape(2 ** k, 2 ** k)
    dim = 2 ** n
    op = np.zeros((dim, dim), dtype=CTYPE)
    bitpos = [n - 1 - a for a in indices]
    for col in range(dim):
        cin = 0
        for j, b in enumerate(bitpos):
            cin |= ((col >> b) & 1) << (k - 1 - j)
        base = col
        for b in bitpos:
            base &= ~(1 << b)
        for r in range(2 ** k):
            row = base
            for j, b in enumerate(bitpos):
                row |= ((r >> (k - 1 - j)) & 1) << b
            op[row, col] += matrix[r, cin]
    return op


def inner(tensor0, tensor1):
    """numpybk.py:125-128: np.vdot (conjugates the first argument, flattens)."""
    return np.vdot(tensor0, tensor1)


def norm(tensor):
    """quantumflow/qubits.py:180-182: |<psi|psi>| (the squared 2-norm)."""
    return np.absolute(inner(tensor, tensor))


def zero_state(n):
    """quantumflow/states.py:186-191: |0...0> as complex128 [2]*n."""
    ket = np.zeros(shape=[2] * n)
    ket[(0,) * n] = 1
    return astensorproduct(ket)


def probabilities(tensor):
    """quantumflow/states.py:113-119: |psi| (np.absolute) then squared."""
    value = np.absolute(tensor)
    return value * value


def expectation(tensor, diag):
    """quantumflow/states.py:131-147 (exact mode, trials=None)."""
    probs = probabilities(tensor)
    diag = astensorproduct(diag)
    return np.sum(np.real(diag) * probs)


def expectation_z(tensor):
    """<Z_i> for every axis i: sum_z |psi_z|^2 (1 - 2 z_i).  Restates
    N calls of states.py:131-147 with diag = +1/-1 on axis i."""
    n = tensor.ndim
    probs = probabilities(tensor)
    out = np.empty(n)
    for i in range(n):
        d = np.ones([2] * n)
        idx = [slice(None)] * n
        idx[i] = 1
        d[tuple(idx)] = -1.0
        out[i] = expectation(tensor, d)
    return out


# --------------------------------------------------------------------------
# Gate matrices (quantumflow/stdgates.py), built with the same numpy
# expressions so entries agree bit for bit with the reference.
# --------------------------------------------------------------------------

def _cis(x):
    # numpybk.py:131-137
    return np.exp(x * 1.0j)


def gate_matrix(name, *params):
    """Return the 2^K x 2^K complex128 matrix of a standard gate.

    Follows the class definitions in quantumflow/stdgates.py (line numbers
    in the table below) with the numpy backend's ccast = identity."""
    p = params
    if name == 'I':                                     # gates.py:33-54
        k = int(p[0]) if p else 1
        u = np.eye(2 ** k)
    elif name == 'X':                                   # stdgates.py:32-48
        u = [[0, 1], [1, 0]]
    elif name == 'Y':                                   # stdgates.py:51-69
        u = np.asarray([[0, -1.0j], [1.0j, 0]])
    elif name == 'Z':                                   # stdgates.py:72-88
        u = [[1, 0], [0, -1.0]]
    elif name == 'H':                                   # stdgates.py:91-109
        u = np.asarray([[1, 1], [1, -1]]) / sqrt(2)
    elif name == 'S':                                   # stdgates.py:112-130
        u = np.asarray([[1.0, 0.0], [0.0, 1.0j]])
    elif name == 'T':                                   # stdgates.py:133-151
        u = [[1.0, 0.0], [0.0, _cis(pi / 4.0)]]
    elif name == 'S_H':                                 # stdgates.py:566-582
        u = np.asarray([[1.0, 0.0], [0.0, -1.0j]])
    elif name == 'T_H':                                 # stdgates.py:585-601
        u = [[1.0, 0.0], [0.0, _cis(-pi / 4.0)]]
    elif name == 'PHASE':                               # stdgates.py:154-176
        u = [[1.0, 0.0], [0.0, _cis(p[0])]]
    elif name == 'RX':                                  # stdgates.py:179-205
        t = p[0]
        u = [[np.cos(t / 2), -1.0j * np.sin(t / 2)],
             [-1.0j * np.sin(t / 2), np.cos(t / 2)]]
    elif name == 'RY':                                  # stdgates.py:208-233
        t = p[0]
        u = [[np.cos(t / 2.0), -np.sin(t / 2.0)],
             [np.sin(t / 2.0), np.cos(t / 2.0)]]
    elif name == 'RZ':                                  # stdgates.py:236-262
        t = p[0]
        u = [[np.exp(-t * 0.5j), 0], [0, np.exp(t * 0.5j)]]
    elif name == 'RN':                                  # stdgates.py:604-636
        t, nx_, ny, nz = p
        c, s = np.cos(t / 2), np.sin(t / 2)
        u = [[c - 1j * s * nz, -1j * s * nx_ - s * ny],
             [-1j * s * nx_ + s * ny, c + 1j * s * nz]]
    elif name in ('TX', 'TY', 'TZ'):                    # stdgates.py:639-758
        t = p[0] % 2
        ct = pi * t
        ph = np.exp(0.5j * ct)
        if name == 'TX':
            u = [[ph * np.cos(ct / 2), ph * -1.0j * np.sin(ct / 2)],
                 [ph * -1.0j * np.sin(ct / 2), ph * np.cos(ct / 2)]]
        elif name == 'TY':
            u = [[ph * np.cos(ct / 2.0), ph * -np.sin(ct / 2.0)],
                 [ph * np.sin(ct / 2.0), ph * np.cos(ct / 2.0)]]
        else:
            u = [[ph * np.exp(-ct * 0.5j), 0], [0, ph * np.exp(ct * 0.5j)]]
    elif name == 'TH':                                  # stdgates.py:761-793
        th = pi * p[0]
        ph = np.exp(0.5j * th)
        a = (ph * 1.0j * np.sin(th / 2)) / sqrt(2)
        u = [[ph * np.cos(th / 2) - a, -a], [-a, ph * np.cos(th / 2) + a]]
    elif name == 'ZYZ':                                 # stdgates.py:796-834
        t0, t1, t2 = (pi * x for x in p)
        u = [[_cis(-0.5 * t2 - 0.5 * t0) * np.cos(0.5 * t1),
              -_cis(-0.5 * t2 + 0.5 * t0) * np.sin(0.5 * t1)],
             [_cis(0.5 * t2 - 0.5 * t0) * np.sin(0.5 * t1),
              _cis(0.5 * t2 + 0.5 * t0) * np.cos(0.5 * t1)]]
    elif name == 'CZ':                                  # stdgates.py:267-288
        u = np.diag([1, 1, 1, -1])
    elif name == 'CNOT':                                # stdgates.py:291-312
        u = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]
    elif name == 'SWAP':                                # stdgates.py:315-338
        u = [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]
    elif name == 'ISWAP':                               # stdgates.py:341-362
        u = np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0],
                      [0, 0, 0, 1]])
    elif name in ('CPHASE00', 'CPHASE01', 'CPHASE10', 'CPHASE'):
        # stdgates.py:365-456
        d = [1.0, 1.0, 1.0, 1.0]
        pos = {'CPHASE00': 0, 'CPHASE01': 1, 'CPHASE10': 2, 'CPHASE': 3}[name]
        d[pos] = np.exp(1j * p[0])
        u = np.diag(d)
    elif name == 'PSWAP':                               # stdgates.py:459-482
        e = np.exp(p[0] * 1.0j)
        u = [[1, 0, 0, 0], [0, 0, e, 0], [0, e, 0, 0], [0, 0, 0, 1]]
    elif name == 'PISWAP':                              # stdgates.py:485-514
        c, s = np.cos(2 * p[0]), np.sin(2 * p[0]) * 1j
        u = [[1, 0, 0, 0], [0, c, s, 0], [0, s, c, 0], [0, 0, 0, 1]]
    elif name in ('XX', 'YY'):                          # stdgates.py:860-913
        th = pi * p[0]
        c, s = np.cos(th / 2), np.sin(th / 2)
        sg = -1.0 if name == 'XX' else 1.0
        u = [[c, 0, 0, sg * 1.0j * s], [0, c, -1.0j * s, 0],
             [0, -1.0j * s, c, 0], [sg * 1.0j * s, 0, 0, c]]
    elif name == 'ZZ':                                  # stdgates.py:944-968
        th = pi * p[0]
        u = np.diag([_cis(-th / 2), _cis(th / 2), _cis(th / 2),
                     _cis(-th / 2)])
    elif name in ('CAN', 'CANONICAL'):                  # stdgates.py:839-857
        xx = gate_matrix('XX', p[0])
        yy = gate_matrix('YY', p[1])
        zz = gate_matrix('ZZ', p[2])
        u = zz @ (yy @ xx)
    elif name == 'EXCH':                                # stdgates.py:971-989
        u = gate_matrix('CAN', p[0], p[0], p[0])
    elif name == 'CCNOT':                               # stdgates.py:519-556
        u = np.eye(8)
        u[[6, 7]] = u[[7, 6]]
    elif name == 'CSWAP':                               # stdgates.py:559-563
        u = np.eye(8)
        u[[5, 6]] = u[[6, 5]]
    else:
        raise ValueError('unknown gate ' + name)
    u = np.asarray(u, dtype=CTYPE)
    return u


def gate_tensor(name, *params):
    """The gate as a [2]*2K tensor (quantumflow/ops.py:101-123)."""
    return astensorproduct(gate_matrix(name, *params))


# --------------------------------------------------------------------------
# Circuits (quantumflow/ops.py:161-166, quantumflow/circuits.py:75-118)
# --------------------------------------------------------------------------

def run_gate(ket, qubits, matrix, gate_qubits):
    """quantumflow/ops.py:161-166: label -> axis lookup, then tensormul."""
    indices = [list(qubits).index(q) for q in gate_qubits]
    k = len(gate_qubits)
    return tensormul(astensorproduct(np.asarray(matrix).reshape(
        [2] * (2 * k))), ket, indices)


def run_circuit(gates, qubits=None, ket=None):
    """quantumflow/circuits.py:87-99: default |0...0> on the sorted qubit set,
    then one tensormul per element.

    ``gates`` is a sequence of (name, params, qubits) or (matrix, qubits)."""
    if qubits is None:
        qbs = sorted(set(q for g in gates for q in g[-1]))
        qubits = tuple(qbs)
    if ket is None:
        ket = zero_state(len(qubits))
    for g in gates:
        if len(g) == 3:
            name, params, gq = g
            m = gate_matrix(name, *params)
        else:
            m, gq = g
        ket = run_gate(ket, qubits, m, gq)
    return ket


def circuit_unitary(gates, qubits):
    """quantumflow/circuits.py:111-118 with ops.py:187-198: identity gate on
    ``qubits`` then ``gate = elem @ gate`` for each element, where gate@gate
    is tensormul of the first gate's tensor into the second."""
    n = len(qubits)
    u = astensorproduct(np.eye(2 ** n))
    for g in gates:
        if len(g) == 3:
            name, params, gq = g
            m = gate_matrix(name, *params)
        else:
            m, gq = g
        k = len(gq)
        indices = [list(qubits).index(q) for q in gq]
        u = tensormul(astensorproduct(np.asarray(m).reshape([2] * (2 * k))),
                      u, indices)
    return u.reshape(2 ** n, 2 ** n)


# --------------------------------------------------------------------------
# DAG layering (quantumflow/dagcircuit.py:97-151)
# --------------------------------------------------------------------------

def dag_depth(gate_qubits, local=True):
    """dagcircuit.py:97-113: longest path in the wire DAG counted in
    operations (in/out nodes excluded).  With local=False, 1-qubit
    operations are dropped first."""
    last = {}
    depth = 0
    for qs in gate_qubits:
        if not local and len(qs) < 2:
            continue
        d = 1 + max((last.get(q, 0) for q in qs), default=0)
        for q in qs:
            last[q] = d
        depth = max(depth, d)
    return depth


def dag_layers(gate_qubits):
    """dagcircuit.py:129-151: element -> its depth (0-based) = 1 + max depth
    of its predecessors on its wires; returns a list of layers of element
    indices."""
    last = {}
    layer_of = []
    for qs in gate_qubits:
        d = max((last.get(q, -1) for q in qs), default=-1) + 1
        for q in qs:
            last[q] = d
        layer_of.append(d)
    nlayers = max(layer_of) + 1 if layer_of else 0
    layers = [[] for _ in range(nlayers)]
    for i, d in enumerate(layer_of):
        layers[d].append(i)
    return layers


# --------------------------------------------------------------------------
# QAOA (quantumflow/qaoa.py, examples/qaoa_maxcut.py)
# --------------------------------------------------------------------------

def qubo_circuit(nodes, edges, steps, beta, gamma, weights=None,
                 node_weights=None):
    """quantumflow/qaoa.py:22-64 as a gate list.

    ``nodes`` in graph.nodes() order, ``edges`` in graph.edges() order,
    ``weights`` per edge (default 1.0), ``node_weights`` dict node->weight
    (an RZ(node_weight) per weighted node, not scaled by gamma, qaoa.py:55-58).
    """
    gates = []
    for q in nodes:
        gates.append(('H', (), (q,)))
    for p in range(steps):
        for e, (q0, q1) in enumerate(edges):
            w = 1.0 if weights is None else weights[e]
            gates.append(('ZZ', (-w * gamma[p] / np.pi,), (q0, q1)))
        for q in nodes:
            if node_weights is not None and q in node_weights:
                gates.append(('RZ', (node_weights[q],), (q,)))
        for q in nodes:
            gates.append(('RX', (beta[p],), (q,)))
    return gates


def graph_cuts_loop(n, edges, weights=None):
    """quantumflow/qaoa.py:68-81 restated literally (O(|E| 2^n) Python loop).
    Node i <-> axis i.  Use only for small n."""
    diag = np.zeros(shape=[2] * n, dtype=np.double)
    for e, (q0, q1) in enumerate(edges):
        w = 1 if weights is None else weights[e]
        for index, _ in np.ndenumerate(diag):
            if index[q0] != index[q1]:
                diag[index] += w
    return diag


def graph_cuts(n, edges, weights=None):
    """Vectorised quantumflow/qaoa.py:68-81: C(z) = sum over edges, in edge
    order, of w_e [z_i != z_j], node i <-> axis i (bit n-1-i of the flat
    index).  Summation order per index equals the reference loop's (adding
    w * 0 leaves a partial sum unchanged), so the table is bitwise identical
    to ``graph_cuts_loop``."""
    idx = np.arange(2 ** n, dtype=np.int64)
    bits = {}

    def bit(q):
        if q not in bits:
            bits[q] = ((idx >> (n - 1 - q)) & 1).astype(np.uint8)
        return bits[q]

    if weights is None:
        count = np.zeros(2 ** n, dtype=np.uint16)
        for (q0, q1) in edges:
            count += bit(q0) ^ bit(q1)
        diag = count.astype(np.double)
    else:
        diag = np.zeros(2 ** n, dtype=np.double)
        for e, (q0, q1) in enumerate(edges):
            diag += weights[e] * (bit(q0) ^ bit(q1))
    return diag.reshape([2] * n)


def maxcut_bruteforce(n, edges, weights=None):
    """Exact integer brute-force MaxCut (max over all 2^n assignments),
    the value examples/qaoa_maxcut.py:57-58 takes as cuts.max()."""
    return float(np.max(graph_cuts(n, edges, weights)))


def qaoa_state(n, edges, beta, gamma, weights=None):
    """Gate-by-gate QAOA state: qubo_circuit then Circuit.run
    (qaoa.py:22-64, circuits.py:87-99).  Nodes are 0..n-1."""
    gates = qubo_circuit(list(range(n)), edges, len(beta), beta, gamma,
                         weights)
    return run_circuit(gates, qubits=tuple(range(n)))


def qaoa_expectation(n, edges, beta, gamma, weights=None):
    """<C> = circ.run().expectation(graph_cuts) (examples/qaoa_maxcut.py:56-59)."""
    ket = qaoa_state(n, edges, beta, gamma, weights)
    return expectation(ket, graph_cuts(n, edges, weights))


def qaoa_state_fast(n, edges, beta, gamma, weights=None, cuts=None):
    """Same state as ``qaoa_state`` computed with the cost layer as one
    diagonal: prod_e ZZ(-w gamma/pi) = exp(i gamma W/2) exp(-i gamma C(z)),
    W = sum of weights (from the ZZ matrix at stdgates.py:944-968).  Mixer
    and H are applied qubit by qubit with tensormul.  Agrees with
    ``qaoa_state`` to rounding; used for sizes where the gate-by-gate
    oracle is slow.  ``cuts``: the graph_cuts table, if already built."""
    if cuts is None:
        cuts = graph_cuts(n, edges, weights)
    cuts = cuts.reshape(-1)
    wsum = float(len(edges) if weights is None else np.sum(weights))
    ket = np.full(2 ** n, 1.0 / np.sqrt(2.0 ** n), dtype=CTYPE)
    ket = ket.reshape([2] * n)
    for p in range(len(beta)):
        phase = np.exp(1j * gamma[p] * (wsum / 2.0 - cuts))
        ket = (ket.reshape(-1) * phase).reshape([2] * n)
        rx = astensorproduct(gate_matrix('RX', beta[p]))
        for q in range(n):
            ket = tensormul(rx, ket, [q])
    return ket


def qaoa_expectation_fast(n, edges, beta, gamma, weights=None, cuts=None):
    if cuts is None:
        cuts = graph_cuts(n, edges, weights)
    ket = qaoa_state_fast(n, edges, beta, gamma, weights, cuts=cuts)
    return expectation(ket, cuts)


def qaoa_value_torch(n, edges, beta, gamma, device='cpu'):
    """<C> as a torch complex128 computation that autograd can differentiate
    (the stored-intermediates gradient the reference gets from TensorFlow
    autograd, examples/qaoa_maxcut.py:53-64).  beta, gamma: float64 tensors
    (requires_grad as needed); cost layer as in ``qaoa_state_fast``, RX on
    every qubit as the elementwise pair update of RX(beta) = [[c, -is],
    [-is, c]] on axis q (state viewed as [2^q, 2, 2^(n-q-1)]).  Unit
    weights."""
    import torch
    cuts = torch.from_numpy(graph_cuts(n, edges).reshape(-1)).to(device)
    wsum = float(len(edges))
    psi = torch.full((2 ** n,), 2.0 ** (-n / 2), dtype=torch.complex128, device=device)
    for l in range(len(beta)):
        psi = psi * torch.exp(1j * gamma[l] * (wsum / 2 - cuts))
        c, s = torch.cos(beta[l] / 2), torch.sin(beta[l] / 2)
        for q in range(n):
            t = psi.reshape(2 ** q, 2, 2 ** (n - q - 1))
            a0, a1 = t[:, 0, :], t[:, 1, :]
            psi = torch.stack([c * a0 - 1j * s * a1, c * a1 - 1j * s * a0], dim=1).reshape(-1)
    return torch.sum(torch.abs(psi) ** 2 * cuts)


def qaoa_grad_fd(n, edges, beta, gamma, h=1e-6, order=2, fast=True,
                 weights=None):
    """Central finite-difference gradient of <C> w.r.t. (beta, gamma).
    order=2: (f(x+h)-f(x-h))/2h; order=4: the 5-point stencil."""
    f = qaoa_expectation_fast if fast else qaoa_expectation
    beta = np.asarray(beta, dtype=float)
    gamma = np.asarray(gamma, dtype=float)
    x = np.concatenate([beta, gamma])
    p = len(beta)

    def ev(xx):
        return f(n, edges, xx[:p], xx[p:], weights)

    g = np.zeros_like(x)
    for i in range(len(x)):
        e = np.zeros_like(x)
        e[i] = h
        if order == 2:
            g[i] = (ev(x + e) - ev(x - e)) / (2 * h)
        else:
            g[i] = (-ev(x + 2 * e) + 8 * ev(x + e) - 8 * ev(x - e)
                    + ev(x - 2 * e)) / (12 * h)
    return g[:p], g[p:]


def qaoa_grad_shift(n, edges, beta, gamma, weights=None):
    """Exact gradient by the parameter-shift rule applied per gate
    occurrence.  RX(b) = exp(-i b X/2) and ZZ(t) = exp(-i pi t Z.Z/2) have
    two-eigenvalue generators, so d<C>/d(angle) of one occurrence equals
    [f(angle + pi/2) - f(angle - pi/2)] / 2 in that occurrence's own angle;
    the chain rule maps ZZ's t = -w gamma/pi (qaoa.py:50-53) to gamma.
    Summed over the occurrences that share a parameter."""
    p = len(beta)
    gates = qubo_circuit(list(range(n)), edges, p, beta, gamma, weights)
    cuts = graph_cuts(n, edges, weights)
    qubits = tuple(range(n))

    def value(gl):
        return expectation(run_circuit(gl, qubits=qubits), cuts)

    gb = np.zeros(p)
    gg = np.zeros(p)
    # locate each occurrence's layer
    pos = n
    for layer in range(p):
        for e in range(len(edges)):
            name, (t,), q = gates[pos]
            w = 1.0 if weights is None else weights[e]
            # ZZ(t) = exp(-i (pi t) Z.Z / 2): angle a = pi t
            plus = list(gates)
            minus = list(gates)
            plus[pos] = (name, (t + 0.5,), q)     # a + pi/2
            minus[pos] = (name, (t - 0.5,), q)    # a - pi/2
            da = (value(plus) - value(minus)) / 2.0
            # a = pi t = -w gamma  ->  da/dgamma = -w
            gg[layer] += da * (-w)
            pos += 1
        for q in range(n):
            name, (b,), qq = gates[pos]
            plus = list(gates)
            minus = list(gates)
            plus[pos] = (name, (b + np.pi / 2,), qq)
            minus[pos] = (name, (b - np.pi / 2,), qq)
            gb[layer] += (value(plus) - value(minus)) / 2.0
            pos += 1
    return gb, gg


def linear_ramp(p, delta=0.75):
    """Linear-ramp QAOA initialisation: gamma_i = (i+1/2)/p * delta,
    beta_i = (1 - (i+1/2)/p) * delta."""
    f = (np.arange(p) + 0.5) / p
    return (1.0 - f) * delta, f * delta


def maxcut_qaoa_gd(n, edges, steps=5, learning_rate=0.01, beta0=None,
                   gamma0=None, max_opt_steps=10000, grad='shift'):
    """examples/qaoa_maxcut.py:37-87 restated: loss = -<C>, plain gradient
    descent with the given learning rate, 10-step blocks, stop when the
    ratio <C>/max C gains less than 1e-4 over a block."""
    cuts = graph_cuts(n, edges)
    maxcut = cuts.max()
    beta = np.array(beta0, dtype=float)
    gamma = np.array(gamma0, dtype=float)
    last_ratio = -1
    ratio = None
    for step in range(0, max_opt_steps, 10):
        for _ in range(10):
            if grad == 'shift':
                gb, gg = qaoa_grad_shift(n, edges, beta, gamma)
            else:
                gb, gg = qaoa_grad_fd(n, edges, beta, gamma)
            beta = beta + learning_rate * gb     # ascend <C> = descend -<C>
            gamma = gamma + learning_rate * gg
        ratio = qaoa_expectation_fast(n, edges, beta, gamma) / maxcut
        if ratio - last_ratio < 1e-4:
            break
        last_ratio = ratio
    return ratio, beta, gamma


# --------------------------------------------------------------------------
# Sampling and measurement (quantumflow/states.py:121-160,
# quantumflow/stdops.py:38-111)
# --------------------------------------------------------------------------

_MASK64 = (1 << 64) - 1


def splitmix64(seed, t):
    """Output t (0-based) of the SplitMix64 generator started at ``seed``
    (Steele, Lea, Flood, OOPSLA 2014; the state advances by the golden-ratio
    increment before each output).  The device sampler keys its draws this
    way; the reference draws through numpy's global generator instead."""
    z = (seed + (t + 1) * 0x9E3779B97F4A7C15) & _MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return z ^ (z >> 31)


def splitmix64_uniform(seed, trials):
    """float64 uniforms in [0, 1): the top 53 bits of each output * 2^-53."""
    return np.array([(splitmix64(seed, t) >> 11) * 2.0 ** -53 for t in range(trials)])


def sample_icdf(tensor, trials, seed):
    """Inverse-CDF draws over |psi|^2 keyed by SplitMix64: draw t is the
    smallest index i with cumsum(p)_i > u_t * sum(p), u_t the t-th uniform.
    Returns (indices, margin) where margin[t] is the distance of u_t * sum(p)
    to the nearest CDF boundary (draws with a margin at rounding level may
    legitimately differ between summation orders).  Restates the sampling
    of states.py:121-129 (np.random.multinomial over probabilities()) as a
    per-draw inverse CDF with a reproducible stream."""
    p = probabilities(np.asarray(tensor)).ravel()
    cdf = np.cumsum(p)
    total = cdf[-1]
    u = splitmix64_uniform(seed, trials) * total
    u = np.minimum(u, np.nextafter(total, 0.0))
    idx = np.searchsorted(cdf, u, side='right')
    lo = np.where(idx > 0, cdf[np.maximum(idx - 1, 0)], 0.0)
    margin = np.minimum(u - lo, cdf[idx] - u)
    return idx.astype(np.int64), margin


def sample_counts_reference(tensor, trials):
    """quantumflow/states.py:121-129: multinomial counts over the
    computational basis from numpy's global generator, shape [2]*N."""
    probs = np.real(probabilities(np.asarray(tensor)))
    res = np.random.multinomial(trials, probs.ravel())
    return res.reshape(probs.shape)


def measure_reference(tensor):
    """quantumflow/states.py:149-160: one basis outcome as N bits."""
    probs = np.real(probabilities(np.asarray(tensor)))
    n = probs.ndim
    indices = np.asarray(list(np.ndindex(*[2] * n)))
    res = np.random.choice(probs.size, p=probs.ravel())
    return indices[res]


def normalize(tensor):
    """quantumflow/states.py:108-111: tensor / sqrt(norm())."""
    return tensor / np.sqrt(norm(tensor))


def measure_qubit(tensor, axis, rand):
    """quantumflow/stdops.py:53-65 with the random roll given: prob_zero is
    P0(q).run(ket).norm(); rand < prob_zero keeps the |0> branch.  Returns
    (post-measurement tensor, bit, prob_zero)."""
    n = tensor.ndim
    p0 = np.array([[1, 0], [0, 0]], dtype=CTYPE)
    p1 = np.array([[0, 0], [0, 1]], dtype=CTYPE)
    ket0 = tensormul(p0.reshape(2, 2), tensor, [axis])
    prob_zero = float(np.real(norm(ket0)))
    if rand < prob_zero:
        return normalize(ket0).reshape([2] * n), 0, prob_zero
    ket1 = tensormul(p1.reshape(2, 2), tensor, [axis])
    return normalize(ket1).reshape([2] * n), 1, prob_zero


def reset_qubits(tensor, axes):
    """quantumflow/stdops.py:100-111: tensormul [[1, 1], [0, 0]] on each
    axis in turn, then normalize."""
    r = np.array([[1, 1], [0, 0]], dtype=CTYPE)
    for a in axes:
        tensor = tensormul(r, tensor, [a])
    return normalize(tensor)


def all_bitstrings(n):
    return np.array(list(itertools.product([0, 1], repeat=n)), dtype=np.int64)


# ---- density matrices -----------------------------------------------------
# A density matrix is a 2^n x 2^n complex128 array, row = ket index (big-
# endian over the qubit axes), the flattening of the reference's rank-2
# QubitVector (quantumflow/states.py:264-344).  Pinned in
# tests/test_oracle_density.py (pure-state consistency, closed forms).

def density_from_state(tensor):
    """|psi><psi| (quantumflow/states.py:162-165)."""
    v = np.asarray(tensor, dtype=CTYPE).reshape(-1)
    return np.outer(v, v.conj())


def density_evolve(rho, matrix, n, indices):
    """U rho U^H for a K-qubit gate on axes ``indices`` (quantumflow/ops.py:
    168-172, whose superoperator ops.py:243-252 is this map), from the full
    2^n x 2^n operator (full_operator)."""
    U = full_operator(matrix, n, indices)
    return U @ np.asarray(rho, dtype=CTYPE) @ U.conj().T


def kraus_evolve(rho, ops, n):
    """sum_k w_k K_k rho K_k^H; ops = [(matrix, indices, weight)]
    (quantumflow/channels.py:74-80)."""
    out = np.zeros((2 ** n, 2 ** n), dtype=CTYPE)
    for matrix, indices, w in ops:
        out += w * density_evolve(rho, matrix, n, indices)
    return out


def measure_density(rho, n, axis, rand):
    """Measure.evolve (quantumflow/stdops.py:67-82) with the uniform draw
    ``rand``: outcome 0 when rand < sum |P0 rho P0|^2 (the projected
    density's norm(), its squared Frobenius norm), the kept projection
    divided by its trace.  Returns (bit, rho)."""
    p0 = np.diag([1.0, 0.0])
    p1 = np.diag([0.0, 1.0])
    r0 = density_evolve(rho, p0, n, [axis])
    if rand < float(np.sum(np.abs(r0) ** 2)):
        return 0, r0 / np.trace(r0)
    r1 = density_evolve(rho, p1, n, [axis])
    return 1, r1 / np.trace(r1)


def density_partial_trace(rho, n, axes):
    """Trace out the qubit axes ``axes`` of an n-qubit density, the others in
    their order (quantumflow/qubits.py:201-229), by an explicit sum over the
    traced indices."""
    keep = [a for a in range(n) if a not in axes]
    t = np.asarray(rho, dtype=CTYPE).reshape([2] * (2 * n))
    m = len(keep)
    out = np.zeros((2 ** m, 2 ** m), dtype=CTYPE)
    for r in itertools.product([0, 1], repeat=m):
        for c in itertools.product([0, 1], repeat=m):
            acc = 0j
            for u in itertools.product([0, 1], repeat=n - m):
                row = [0] * n
                col = [0] * n
                for j, a in enumerate(keep):
                    row[a], col[a] = r[j], c[j]
                for j, a in enumerate(axes):
                    row[a] = col[a] = u[j]
                acc += t[tuple(row + col)]
            out[int(''.join(map(str, r)) or '0', 2), int(''.join(map(str, c)) or '0', 2)] = acc
    return out

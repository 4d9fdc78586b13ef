"""QAOA for MaxCut (quantumflow/qaoa.py:1-81, examples/qaoa_maxcut.py:37-87).

``qubo_circuit`` builds the reference's gate-level circuit (H on every node,
then per step ZZ(-w gamma/pi) per edge, RZ(node weight) per weighted node and
RX(beta) per node) -- the plain per-gate path.  ``graph_cuts`` computes the
cost diagonal on the GPU.  The batched, fused and differentiable engine is
in quantumflow_amd.qaoa_engine.
"""

from typing import Sequence

import networkx as nx
import numpy as np
import torch

from . import backend as bk
from . import _native
from .circuits import Circuit
from .stdgates import H, RX, RZ, ZZ

__all__ = ['qubo_circuit', 'graph_cuts', 'graph_edges']


def qubo_circuit(graph: nx.Graph, steps: int, beta: Sequence, gamma: Sequence) -> Circuit:
    """QAOA circuit for the Ising/QUBO problem of ``graph`` (qaoa.py:22-64)."""
    qubits = list(graph.nodes())
    circ = Circuit()
    for q0 in qubits:
        circ += H(q0)
    for p in range(0, steps):
        for q0, q1 in graph.edges():
            weight = graph[q0][q1].get('weight', 1.0)
            circ += ZZ(-weight * gamma[p] / np.pi, q0, q1)
        for q0 in qubits:
            node_weight = graph.nodes[q0].get('weight', None)
            if node_weight is not None:
                circ += RZ(node_weight, q0)
        for q0 in qubits:
            circ += RX(beta[p], q0)
    return circ


def graph_edges(graph: nx.Graph):
    """(n, edges int32 [m,2], weights float64 [m] or None) in graph.edges()
    order; node labels must be 0..n-1 (qaoa.py:73-79 indexes axes by label)."""
    n = len(graph)
    edges = np.array([[u, v] for u, v in graph.edges()], dtype=np.int32).reshape(-1, 2)
    if edges.size and (edges.min() < 0 or edges.max() >= n):
        raise ValueError('graph_cuts: node labels must be 0..n-1')
    ws = [graph[u][v].get('weight', None) for u, v in graph.edges()]
    weights = None
    if any(w is not None for w in ws):
        weights = np.array([1.0 if w is None else float(w) for w in ws], dtype=np.float64)
    return n, edges, weights


def graph_cuts(graph: nx.Graph) -> bk.BKTensor:
    """Cut value of every assignment as a float64 [2]*N device tensor,
    computed by one kernel (bitwise equal to qaoa.py:68-81)."""
    n, edges, weights = graph_edges(graph)
    dev = bk.DEVICE
    out = torch.empty(2 ** n, dtype=torch.float64, device=dev)
    e = torch.from_numpy(edges.copy()).to(dev) if len(edges) else \
        torch.zeros(2, dtype=torch.int32, device=dev)
    w = torch.from_numpy(weights).to(dev) if weights is not None else None
    _native.call('qfx_cut_table', bk.stream_handle(), _native.ctypes.c_void_p(out.data_ptr()), n,
                 _native.ctypes.c_void_p(e.data_ptr()),
                 None if w is None else _native.ctypes.c_void_p(w.data_ptr()), len(edges))
    return out.view([2] * n)


def plain_value(graph: nx.Graph, beta: Sequence, gamma: Sequence):
    """<C> by the reference's chain, one kernel launch per gate:
    qubo_circuit -> Circuit.run -> State.expectation(graph_cuts)
    (examples/qaoa_maxcut.py:56-59).  Returns a 0-dim device tensor."""
    from .states import zero_state
    circ = qubo_circuit(graph, len(beta), beta, gamma)
    ket = circ.run(zero_state(list(graph.nodes())), fused=False)
    return ket.expectation(graph_cuts(graph))


def plain_value_and_grad(graph: nx.Graph, beta: Sequence, gamma: Sequence):
    """<C> and d<C>/d(beta, gamma) on the plain per-gate path: the forward
    gate by gate, then an adjoint sweep gate by gate (psi and lambda carried
    back by each gate's inverse; every RX and ZZ occurrence adds its
    <lambda|dG|psi> term).  Returns (value, grad [2p]) as device tensors."""
    from .fusion import apply_gates_inplace
    from . import stdgates
    p = len(beta)
    n = len(graph)
    qubits = tuple(range(n))
    circ = qubo_circuit(graph, p, beta, gamma)
    gates = circ.elements
    psi = bk.zero_state_tensor(n).reshape(-1).clone()
    apply_gates_inplace(gates, qubits, psi)
    cuts = graph_cuts(graph).reshape(-1)
    value = bk.expectation(cuts, psi)
    lam = (psi * cuts).contiguous()
    tmp = torch.empty_like(psi)
    grad = torch.zeros(2 * p, dtype=torch.float64, device=psi.device)
    zz_kind = bk.classify(np.diag([1.0, -1.0, -1.0, 1.0]))
    # layer of each parametrised gate occurrence, in circuit order
    layer_of = []
    for l in range(p):
        layer_of += [l] * (graph.number_of_edges() + n)
    pos = len(gates) - 1
    occ = len(layer_of) - 1
    while pos >= 0:
        g = gates[pos]
        bits = [n - 1 - qubits.index(q) for q in g.qubits]
        if isinstance(g, (stdgates.RX, stdgates.ZZ)):
            l = layer_of[occ]
            occ -= 1
            if isinstance(g, stdgates.RX):
                # d/dbeta RX(beta) = (-i X/2) RX: term Im <lam|X|psi>
                bk.apply_gate(tmp, psi, n, bits, stdgates.X().kind)
                grad[l] += bk.inner(lam, tmp).imag
            else:
                # ZZ(t), t = -w gamma/pi: d/dgamma = (i w / 2) Z.Z ZZ(t)
                w = float(graph[g.qubits[0]][g.qubits[1]].get('weight', 1.0))
                bk.apply_gate(tmp, psi, n, bits, zz_kind)
                grad[p + l] += -w * bk.inner(lam, tmp).imag
        inv = g.H.kind
        bk.apply_gate(psi, psi, n, bits, inv, inplace_sparse=True)
        bk.apply_gate(lam, lam, n, bits, inv, inplace_sparse=True)
        pos -= 1
    return value, grad


def autograd_value_and_grad(graph: nx.Graph, beta: Sequence, gamma: Sequence):
    """<C> and its gradient the way the reference gets it
    (examples/qaoa_maxcut.py:53-64, TensorFlow autograd over the gate-level
    circuit): every gate of ``qubo_circuit`` after the initial layer of H is
    a torch complex128 operation on the device -- ZZ(-w gamma/pi) on (u, v)
    multiplies by exp(i w gamma z_u z_v / 2), RX(beta) updates the amplitude
    pairs of its axis -- and autograd keeps what each gate's backward needs
    (the state entering it), so memory grows with the gate count.  The
    stored-intermediates comparator of the adjoint engine; unit or 'weight'
    edge weights, no node weights.  Returns (value, grad [2p]) as device
    tensors."""
    p = len(beta)
    n = len(graph)
    dev = bk.DEVICE
    b = torch.tensor(np.asarray(beta, dtype=np.float64), device=dev, requires_grad=True)
    g = torch.tensor(np.asarray(gamma, dtype=np.float64), device=dev, requires_grad=True)
    cuts = graph_cuts(graph).reshape(-1)
    idx = torch.arange(2 ** n, device=dev)
    psi = torch.full((2 ** n,), 2.0 ** (-n / 2), dtype=torch.complex128, device=dev)
    for layer in range(p):
        for u, v in graph.edges():
            w = float(graph[u][v].get('weight', 1.0))
            zz = 1.0 - 2.0 * (((idx >> (n - 1 - u)) ^ (idx >> (n - 1 - v))) & 1).to(torch.float64)
            psi = psi * torch.exp(1j * (0.5 * w) * g[layer] * zz)
        c, s = torch.cos(b[layer] / 2), torch.sin(b[layer] / 2)
        for q in range(n):
            t = psi.reshape(2 ** q, 2, 2 ** (n - q - 1))
            a0, a1 = t[:, 0, :], t[:, 1, :]
            psi = torch.stack([c * a0 - 1j * s * a1, c * a1 - 1j * s * a0], dim=1).reshape(-1)
    value = torch.sum((psi.real ** 2 + psi.imag ** 2) * cuts)
    value.backward()
    return value.detach(), torch.cat([b.grad, g.grad])

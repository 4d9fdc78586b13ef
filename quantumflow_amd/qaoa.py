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

"""Standard graph sets (quantumflow/datasets/__init__.py:23-37).

``load_stdgraphs(size)`` returns the 100 Erdos-Renyi graphs (edge
probability 1/2) on ``size`` nodes, 6 <= size <= 32, that QuantumFlow ships
as graph6 files; the files under ``data/`` are those data files.
"""

import os

import networkx as nx

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data')


def load_stdgraphs(size: int):
    """100 graphs on ``size`` nodes; ValueError outside 6..32 (as the
    reference)."""
    if size < 6 or size > 32:
        raise ValueError('Size out of range.')
    path = os.path.join(_DATA, 'graph{}er100.g6'.format(size))
    with open(path, 'rb') as f:
        return nx.read_graph6(f)

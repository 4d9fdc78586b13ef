"""Writes stdgraph_maxcut.json: exact MaxCut of every graph in the standard
graph sets for 6..16 nodes (quantumflow/datasets/data/graph{n}er100.g6,
shipped as quantumflow_amd/datasets/data), by brute force over all 2^n
two-colourings with numpy bit arithmetic -- independent of the oracle's
graph_cuts and of the GPU path, both of which the tests check against it.

    python tests/golden/make_stdgraph_maxcut.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def brute_force(graph):
    n = graph.number_of_nodes()
    z = np.arange(2 ** n, dtype=np.int64)
    cut = np.zeros(2 ** n, dtype=np.int64)
    for u, v in graph.edges():
        cut += ((z >> u) ^ (z >> v)) & 1
    best = int(cut.max())
    return best, int(np.count_nonzero(cut == best))


def main():
    from quantumflow_amd.datasets import load_stdgraphs
    out = {"source": "quantumflow/datasets/data/graph{n}er100.g6 (100 ER(1/2) graphs per n)",
           "method": "brute force over all 2^n colourings, unit weights",
           "sizes": {}}
    for n in range(6, 17):
        rows = [brute_force(g) for g in load_stdgraphs(n)]
        out["sizes"][str(n)] = {"maxcut": [r[0] for r in rows],
                                "optimal_colourings": [r[1] for r in rows]}
    with open(os.path.join(HERE, "stdgraph_maxcut.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()

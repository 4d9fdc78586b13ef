"""The standard graph sets (quantumflow/datasets/__init__.py:23-37) and the
oracle's cut tables against the brute-force MaxCut fixture
(tests/golden/stdgraph_maxcut.json, written by make_stdgraph_maxcut.py)."""
import json
import os

import numpy as np
import pytest

from oracle import qf_oracle as O
from quantumflow_amd.datasets import load_stdgraphs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "stdgraph_maxcut.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)["sizes"]


def test_load_stdgraphs_shape():
    for n in (6, 10, 20, 32):
        graphs = load_stdgraphs(n)
        assert len(graphs) == 100
        assert all(sorted(g.nodes()) == list(range(n)) for g in graphs)
    with pytest.raises(ValueError):
        load_stdgraphs(5)
    with pytest.raises(ValueError):
        load_stdgraphs(33)


@pytest.mark.parametrize("n", [6, 8, 10, 12])
def test_oracle_maxcut_matches_fixture(n):
    want = golden()[str(n)]
    for g, best, count in zip(load_stdgraphs(n), want["maxcut"], want["optimal_colourings"]):
        cuts = O.graph_cuts(n, list(g.edges())).reshape(-1)
        assert cuts.max() == best
        assert np.count_nonzero(cuts == best) == count


def test_graph_cuts_loop_bitwise_on_fixture_graphs():
    for g in load_stdgraphs(6)[:10]:
        e = list(g.edges())
        assert np.array_equal(O.graph_cuts_loop(6, e), O.graph_cuts(6, e))

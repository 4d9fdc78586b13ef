"""Autotuner cache and defaults (CPU), and one tuning run of the batched
QAOA engine on the GPU."""
import numpy as np
import pytest

from quantumflow_amd import autotune, config


@pytest.fixture
def cache(tmp_path, monkeypatch):
    monkeypatch.setattr(config, 'AUTOTUNE_CACHE', str(tmp_path / 'autotune.json'))
    monkeypatch.setattr(autotune, '_cache', None)
    yield tmp_path / 'autotune.json'
    autotune._cache = None


def test_qaoa_defaults_and_store(cache):
    assert autotune.choose_qaoa(0) == {'tile_bits': 13, 'block': 512}
    assert autotune.choose_qaoa(1) == {'tile_bits': 12, 'block': 512}
    autotune.store_qaoa(1, {'tile_bits': 11, 'block': 256}, 3.5)
    assert autotune.choose_qaoa(1) == {'tile_bits': 11, 'block': 256}
    assert autotune.choose_qaoa(0) == {'tile_bits': 13, 'block': 512}
    autotune._cache = None                      # reread from disk
    assert autotune.choose_qaoa(1) == {'tile_bits': 11, 'block': 256}
    assert cache.exists()


def test_circuit_defaults(cache):
    cfg = autotune.choose(28)
    assert cfg['tile_bits'] == 12 and cfg['block'] == 256
    small = autotune.choose(9)
    assert small['tile_bits'] == 9 and (1 << 9) >= small['block']
    # 16-amplitude contiguous runs once the state outgrows the 256 MiB cache
    assert [autotune.choose(n)['low_bits'] for n in (9, 16, 24, 25, 28)] == [3, 3, 3, 4, 4]


@pytest.mark.gpu
def test_tune_qaoa_picks_fastest_and_batches_use_it(cache):
    import networkx as nx
    from quantumflow_amd import qaoa_engine as E
    graphs = [nx.convert_node_labels_to_integers(nx.random_regular_graph(3, n, seed=n))
              for n in (14, 16, 16, 18)]
    cands = [dict(tile_bits=12, block=512), dict(tile_bits=11, block=256)]
    best, results = autotune.tune_qaoa(graphs, 3, E.MODE_BACKWARD, reps=2, candidates=cands)
    assert best in cands
    assert results[tuple(sorted(best.items()))] == min(results.values())
    batch = E.QAOABatch(graphs, 3, mode=E.MODE_BACKWARD)
    assert (batch.T, batch.block) == (best['tile_bits'], best['block'])
    # every candidate computes the same values and gradients
    rng = np.random.default_rng(1)
    beta, gamma = rng.uniform(0, 1, (4, 3)), rng.uniform(0, 2, (4, 3))
    outs = [E.QAOABatch(graphs, 3, mode=E.MODE_BACKWARD, **c).evaluate(beta, gamma)
            for c in cands]
    for a, b in zip(outs[0], outs[1]):
        assert np.max(np.abs(a.cpu().numpy() - b.cpu().numpy())) < 1e-11

"""Pins for the MLP oracle: backward vs central finite differences of the
loss, Adam's first step (= -lr * sign(g) up to eps), and the proposer's
graph features against hand-counted values."""
import networkx as nx
import numpy as np

from oracle import mlp_oracle as M


def rand_params(sizes, rng):
    return [(rng.normal(size=(o, i)) / np.sqrt(i), rng.normal(size=o) * 0.1)
            for i, o in zip(sizes[:-1], sizes[1:])]


def test_backward_vs_finite_differences():
    rng = np.random.default_rng(0)
    params = rand_params([5, 7, 6, 3], rng)
    x = rng.normal(size=(4, 5))
    t = rng.normal(size=(4, 3))
    loss, grads = M.loss_and_grads(params, x, t)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            flat = arr.reshape(-1)
            gflat = g.reshape(-1)
            for j in range(0, flat.size, max(1, flat.size // 7)):
                old = flat[j]
                flat[j] = old + h
                lp, _ = M.loss_and_grads(params, x, t)
                flat[j] = old - h
                lm, _ = M.loss_and_grads(params, x, t)
                flat[j] = old
                assert abs((lp - lm) / (2 * h) - gflat[j]) < 1e-7


def test_adam_first_step_closed_form():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p1, m, v = M.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # bias-corrected first step: m_hat = g, v_hat = g^2 -> step = lr g/(|g|+eps)
    assert np.allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8), atol=1e-15)


def test_features_hand_counted():
    from quantumflow_amd.proposer import graph_features, FEATURES
    g = nx.Graph([(0, 1), (1, 2), (2, 0), (2, 3)])          # triangle + pendant
    f = graph_features(g, normalise=False)
    assert len(f) == FEATURES == 19
    assert f[0] == 4 and f[1] == 4                          # n, |E|
    hist = f[2:18]
    assert hist[1] == 1 and hist[2] == 2 and hist[3] == 1 and hist.sum() == 4
    assert f[18] == 1                                        # one triangle

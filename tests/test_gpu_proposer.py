"""GPU checks of the angle proposer: its hand-written fp32 kernels against
the float64 MLP oracle (forward, gradients, Adam), training reduces the
loss, and a trained proposer starts the QAOA optimizer better than the
linear-ramp fallback on fresh graphs."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    import quantumflow_amd.proposer as P
    import quantumflow_amd.backend as bk
    assert bk.gpu_available()
    return P


def host_params(prop):
    flat = prop.params.double().cpu().numpy()
    out = []
    for wo, bo, fi, fo in prop.layout:
        out.append((flat[wo:wo + fi * fo].reshape(fo, fi), flat[bo:bo + fo]))
    return out


def test_forward_and_gradients_vs_oracle(P):
    import torch
    prop = P.AngleProposer(p=3, hidden=(40, 33), seed=3)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(70, P.FEATURES)).astype(np.float32)
    t = rng.normal(size=(70, 6)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(t).cuda()
    params = host_params(prop)
    acts = M.forward(params, x.astype(np.float64))
    y = prop.forward(xd).double().cpu().numpy()
    assert np.max(np.abs(y - acts[-1])) <= 1e-5 * (1 + np.max(np.abs(acts[-1])))
    loss = prop.loss_and_grad(xd, td).item()
    ref_loss, ref_grads = M.loss_and_grads(params, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-5 * (1 + ref_loss)
    flat = prop.grads.double().cpu().numpy()
    for (wo, bo, fi, fo), (dW, db) in zip(prop.layout, ref_grads):
        gw = flat[wo:wo + fi * fo].reshape(fo, fi)
        gb = flat[bo:bo + fo]
        scale = 1e-4 * (1 + np.max(np.abs(dW)))
        assert np.max(np.abs(gw - dW)) <= scale
        assert np.max(np.abs(gb - db)) <= 1e-4 * (1 + np.max(np.abs(db)))


def test_adam_vs_oracle(P):
    prop = P.AngleProposer(p=2, hidden=(16,), seed=4)
    rng = np.random.default_rng(1)
    import torch
    g = rng.normal(size=prop.nparams).astype(np.float32)
    p0 = prop.params.double().cpu().numpy()
    prop.grads.copy_(torch.from_numpy(g).cuda())
    prop.adam_step(lr=0.01, weight_decay=0.001)
    prop.adam_step(lr=0.01, weight_decay=0.001)
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    p = p0
    for step in (1, 2):
        p, m, v = M.adam(p, g.astype(np.float64), m, v, step, lr=0.01, wd=0.001)
    assert np.max(np.abs(prop.params.double().cpu().numpy() - p)) <= 1e-6


def test_untrained_proposer_is_linear_ramp(P):
    import networkx as nx
    prop = P.AngleProposer(p=4, seed=0)
    prop.params.view(-1)[:prop.layout[-1][0]].zero_()     # silence the hidden path
    b, g = prop.propose([nx.cycle_graph(5)])
    rb, rg = P.linear_ramp(4)
    assert np.allclose(b[0], rb, atol=1e-6) and np.allclose(g[0], rg, atol=1e-6)


def test_training_reduces_loss(P):
    rng = np.random.default_rng(2)
    X = rng.normal(size=(512, P.FEATURES)).astype(np.float32)
    W = rng.normal(size=(P.FEATURES, 8)).astype(np.float32) * 0.3
    T = np.tanh(X @ W).astype(np.float32)
    prop = P.AngleProposer(p=4, hidden=(64, 64), seed=5)
    hist = prop.fit(X, T, epochs=30, batch=128, lr=3e-3)
    assert hist[-1] < 0.2 * hist[0]


def test_trained_proposer_beats_linear_ramp(P):
    """Small end-to-end version of the learned warm start: optimise 400
    graphs (n 8..12) from the linear ramp, train on (features, optimised
    angles), then compare the starting ratio <C>/maxcut on 120 fresh graphs."""
    p = 3
    train = P.random_graphs(400, 8, 12, seed=10)
    rb, rg = P.linear_ramp(p)
    b0 = np.tile(rb, (len(train), 1))
    g0 = np.tile(rg, (len(train), 1))
    bopt, gopt, ratio, _ = P.optimise_graphs(train, p, b0, g0, steps=120, lr=0.02)
    assert np.mean(ratio) > 0.7
    X = np.stack([P.graph_features(g) for g in train]).astype(np.float32)
    T = np.concatenate([bopt, gopt], axis=1).astype(np.float32)
    prop = P.AngleProposer(p=p, hidden=(128, 128, 128), seed=6)
    prop.fit(X, T, epochs=150, batch=64, lr=1e-3)
    test = P.random_graphs(120, 8, 12, seed=11)
    pb, pg = prop.propose(test)
    r_mlp = P.expectation_ratios(test, p, pb, pg)
    r_ramp = P.expectation_ratios(test, p, np.tile(rb, (120, 1)), np.tile(rg, (120, 1)))
    print("start ratio: proposer %.4f  linear ramp %.4f" % (r_mlp.mean(), r_ramp.mean()))
    assert r_mlp.mean() > r_ramp.mean() + 0.01

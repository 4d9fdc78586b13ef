"""GPU parity of the batched QAOA engine against the oracle.

Values: |d<C>| <= 1e-10 (relative to |E|-sized sums, complex128).
Gradients: <= 1e-10 against the exact parameter-shift rule (small n) and a
CPU torch complex128 autograd restatement; against central finite
differences with h=1e-6 to 1e-7 (rounding floor ~ eps |<C>| / h) and the
4th-order stencil at h=1e-3 to 1e-9.
"""
import networkx as nx
import numpy as np
import pytest

from oracle import qf_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import quantumflow_amd.qaoa_engine as E
    import quantumflow_amd.backend as bk
    assert bk.gpu_available()
    return E


def er(n, pe, seed):
    return nx.gnp_random_graph(n, pe, seed=seed)


def test_forward_values_mixed_sizes(E):
    rng = np.random.default_rng(1)
    graphs = [er(n, 0.5, 100 + n) for n in (3, 6, 10, 12, 13, 14, 16, 17, 19)]
    p = 3
    beta = rng.uniform(0, np.pi / 2, size=(len(graphs), p))
    gamma = rng.uniform(0, np.pi, size=(len(graphs), p))
    print("forward batch", [len(g) for g in graphs], flush=True)
    batch = E.QAOABatch(graphs, p, mode=E.MODE_FORWARD)
    val, _ = batch.evaluate(beta, gamma)
    batch.check()
    val = val.cpu().numpy()
    for i, g in enumerate(graphs):
        n = len(g)
        ref = O.qaoa_expectation_fast(n, list(g.edges()), beta[i], gamma[i])
        assert abs(val[i] - ref) <= 1e-10, (n, val[i], ref)
    for i in (0, 2):
        g = graphs[i]
        ref = O.qaoa_expectation(len(g), list(g.edges()), beta[i], gamma[i])   # gate by gate
        assert abs(val[i] - ref) <= 1e-10


def test_backward_vs_parameter_shift(E):
    rng = np.random.default_rng(2)
    graphs = [er(6, 0.5, 7), er(8, 0.4, 8), nx.cycle_graph(7)]
    p = 2
    beta = rng.uniform(0, np.pi / 2, size=(3, p))
    gamma = rng.uniform(0, np.pi, size=(3, p))
    batch = E.QAOABatch(graphs, p, mode=E.MODE_BACKWARD)
    val, grad = batch.evaluate(beta, gamma)
    batch.check()
    val, grad = val.cpu().numpy(), grad.cpu().numpy()
    for i, g in enumerate(graphs):
        n, edges = len(g), list(g.edges())
        assert abs(val[i] - O.qaoa_expectation(n, edges, beta[i], gamma[i])) <= 1e-12
        sb, sg = O.qaoa_grad_shift(n, edges, beta[i], gamma[i])
        assert np.max(np.abs(grad[i, :p] - sb)) <= 1e-10
        assert np.max(np.abs(grad[i, p:] - sg)) <= 1e-10


def _torch_value(n, edges, beta, gamma):
    """CPU torch complex128 restatement of the QAOA value, for autograd."""
    import torch
    cuts = torch.from_numpy(O.graph_cuts(n, edges).reshape(-1))
    W = float(len(edges))
    psi = torch.full((2 ** n,), 2.0 ** (-n / 2), dtype=torch.complex128)
    for l in range(len(beta)):
        psi = psi * torch.exp(1j * gamma[l] * (W / 2 - cuts))
        c, s = torch.cos(beta[l] / 2), torch.sin(beta[l] / 2)
        m = torch.stack([torch.stack([c + 0j, -1j * s]), torch.stack([-1j * s, c + 0j])])
        t = psi.reshape([2] * n)
        for q in range(n):
            t = torch.movedim(torch.tensordot(m, t, dims=([1], [q])), 0, q)
        psi = t.reshape(-1)
    return torch.sum(torch.abs(psi) ** 2 * cuts)


def test_backward_multipass_vs_autograd_and_fd(E):
    import torch
    rng = np.random.default_rng(3)
    graphs = [er(14, 0.4, 31), nx.random_regular_graph(3, 16, seed=5), er(13, 0.5, 33)]
    p = 3
    beta = rng.uniform(0, np.pi / 2, size=(3, p))
    gamma = rng.uniform(0, np.pi, size=(3, p))
    batch = E.QAOABatch(graphs, p, mode=E.MODE_BACKWARD)
    val, grad = batch.evaluate(beta, gamma)
    batch.check()
    val, grad = val.cpu().numpy(), grad.cpu().numpy()
    for i, g in enumerate(graphs):
        n, edges = len(g), list(g.edges())
        b = torch.tensor(beta[i], requires_grad=True)
        gm = torch.tensor(gamma[i], requires_grad=True)
        v = _torch_value(n, edges, b, gm)
        v.backward()
        assert abs(val[i] - v.item()) <= 1e-10
        assert np.max(np.abs(grad[i, :p] - b.grad.numpy())) <= 1e-10
        assert np.max(np.abs(grad[i, p:] - gm.grad.numpy())) <= 1e-10
        fb, fg = O.qaoa_grad_fd(n, edges, beta[i], gamma[i], h=1e-6)
        assert np.max(np.abs(grad[i, :p] - fb)) <= 1e-7
        assert np.max(np.abs(grad[i, p:] - fg)) <= 1e-7
        fb4, fg4 = O.qaoa_grad_fd(n, edges, beta[i], gamma[i], h=1e-3, order=4)
        assert np.max(np.abs(grad[i, :p] - fb4)) <= 1e-9
        assert np.max(np.abs(grad[i, p:] - fg4)) <= 1e-9


def test_weighted_graph(E):
    rng = np.random.default_rng(4)
    g = er(9, 0.5, 44)
    for u, v in g.edges():
        g[u][v]['weight'] = float(rng.uniform(0.1, 2.0))
    edges = list(g.edges())
    w = [g[u][v]['weight'] for u, v in edges]
    p = 2
    beta = rng.uniform(0, 1, size=(1, p))
    gamma = rng.uniform(0, 1, size=(1, p))
    batch = E.QAOABatch([g], p, mode=E.MODE_BACKWARD)
    val, grad = batch.evaluate(beta, gamma)
    val, grad = val.cpu().numpy(), grad.cpu().numpy()
    assert abs(val[0] - O.qaoa_expectation(9, edges, beta[0], gamma[0], w)) <= 1e-10
    sb, sg = O.qaoa_grad_shift(9, edges, beta[0], gamma[0], w)
    assert np.max(np.abs(grad[0] - np.concatenate([sb, sg]))) <= 1e-10


@pytest.mark.parametrize("mode", [0, 1])
def test_fused_cost_stages_match_separate_stages(E, mode):
    """The default program (forward: cost layers fused into mixer chunks,
    RX bits applied in registers during tile loads and stores) against the
    same program with separate stages, on multi-pass instances at several
    tile widths and workgroup sizes, plus a weighted multi-pass instance
    (whose fused stages run as their two parts) against the oracle."""
    import torch
    rng = np.random.default_rng(8 + mode)
    graphs = [er(n, 0.5, 80 + n) for n in (7, 11, 14, 17)]
    wg = er(12, 0.5, 99)
    for u, v in wg.edges():
        wg[u][v]['weight'] = float(rng.uniform(0.1, 2.0))
    graphs.append(wg)
    p = 3
    beta = rng.uniform(0, np.pi / 2, size=(len(graphs), p))
    gamma = rng.uniform(0, np.pi, size=(len(graphs), p))
    configs = [(None, 512), (8, 512)] if mode == 1 else [(None, 512), (13, 256), (10, 256), (8, 512)]
    for T, block in configs:
        print("fused vs separate, mode", mode, "tile bits", T, "block", block, flush=True)
        va, ga = E.QAOABatch(graphs, p, mode=mode, tile_bits=T, block=block).evaluate(beta, gamma)
        vb, gb = E.QAOABatch(graphs, p, mode=mode, tile_bits=T, block=block,
                             fuse=False).evaluate(beta, gamma)
        assert torch.max(torch.abs(va - vb)).item() <= 1e-12
        if mode == 1:
            assert torch.max(torch.abs(ga - gb)).item() <= 1e-12
    edges = list(wg.edges())
    w = [wg[u][v]['weight'] for u, v in edges]
    i = len(graphs) - 1
    assert abs(va[i].item() - O.qaoa_expectation(12, edges, beta[i], gamma[i], w)) <= 1e-10
    if mode == 1:
        sb, sg = O.qaoa_grad_shift(12, edges, beta[i], gamma[i], w)
        assert np.max(np.abs(ga[i].cpu().numpy() - np.concatenate([sb, sg]))) <= 1e-10


def test_graph_replay_and_determinism(E):
    import torch
    rng = np.random.default_rng(5)
    graphs = [er(n, 0.5, 50 + n) for n in (8, 11, 15, 18)]
    p = 2
    batch = E.QAOABatch(graphs, p, mode=E.MODE_BACKWARD)
    beta = rng.uniform(0, 1, size=(4, p))
    gamma = rng.uniform(0, 1, size=(4, p))
    v0, g0 = batch.evaluate(beta, gamma)
    v0, g0 = v0.clone(), g0.clone()
    v1, g1 = batch.evaluate(beta, gamma)
    assert torch.equal(v0, v1) and torch.equal(g0, g1)          # bitwise repeatable
    for rep in range(3):
        vg, gg = batch.evaluate(beta, gamma, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(vg, v0) and torch.equal(gg, g0), rep
        batch.evaluate(beta * 0.5, gamma, graph=False)          # eager work between replays
        torch.cuda.synchronize()
    batch.check()


def test_batch_equals_singletons(E):
    import torch
    rng = np.random.default_rng(6)
    graphs = [er(n, 0.45, 60 + n) for n in (10, 14, 12)]
    p = 2
    beta = rng.uniform(0, 1, size=(3, p))
    gamma = rng.uniform(0, 1, size=(3, p))
    vb, gb = E.QAOABatch(graphs, p).evaluate(beta, gamma)
    for i, g in enumerate(graphs):
        vs, gs = E.QAOABatch([g], p).evaluate(beta[i:i + 1], gamma[i:i + 1])
        assert abs(vb[i].item() - vs[0].item()) <= 1e-12
        assert torch.max(torch.abs(gb[i] - gs[0])).item() <= 1e-12


def test_autograd_function(E):
    import torch
    graphs = [er(9, 0.5, 71), er(12, 0.5, 72)]
    p = 2
    batch = E.QAOABatch(graphs, p)
    beta = torch.full((2, p), 0.3, dtype=torch.float64, requires_grad=True)
    gamma = torch.full((2, p), 0.4, dtype=torch.float64, requires_grad=True)
    val = E.qaoa_expectation(beta, gamma, batch)
    (val * torch.tensor([1.0, 2.0], dtype=torch.float64, device=val.device)).sum().backward()
    _, grad = batch.evaluate(beta.detach(), gamma.detach())
    g = grad.cpu()
    assert torch.allclose(beta.grad.cpu(), torch.stack([g[0, :p], 2 * g[1, :p]]), atol=1e-14)
    assert torch.allclose(gamma.grad.cpu(), torch.stack([g[0, p:], 2 * g[1, p:]]), atol=1e-14)


def test_optimizer_star_kat(E):
    # tests/test_qaoa_maxcut.py:12-21: 4-node star, p=5, ratio > 0.95
    g = nx.Graph([(0, 1), (1, 2), (1, 3)])
    batch = E.QAOABatch([g], 5)
    rng = np.random.default_rng(7)
    b0 = rng.normal(0.5, 0.01, size=(1, 5))
    g0 = rng.normal(0.5, 0.01, size=(1, 5))
    ratio, beta, gamma, steps = E.optimize(batch, b0, g0, learning_rate=0.01, max_steps=4000)
    assert ratio[0].item() > 0.95


def test_plain_path_matches_engine(E):
    from quantumflow_amd import qaoa as Q
    rng = np.random.default_rng(8)
    graphs = [er(7, 0.5, 81), er(11, 0.5, 82)]
    p = 2
    beta = rng.uniform(0, 1, size=(2, p))
    gamma = rng.uniform(0, 1, size=(2, p))
    vb, gb = E.QAOABatch(graphs, p).evaluate(beta, gamma)
    for i, g in enumerate(graphs):
        v1 = Q.plain_value(g, beta[i], gamma[i]).item()
        v2, g2 = Q.plain_value_and_grad(g, beta[i], gamma[i])
        assert abs(v1 - vb[i].item()) <= 1e-10
        assert abs(v2.item() - vb[i].item()) <= 1e-10
        assert np.max(np.abs(g2.cpu().numpy() - gb[i].cpu().numpy())) <= 1e-10


def test_engine_maxcut_on_standard_graph_sets(E):
    """Exact MaxCut from the engine's cut tables equals the brute-force
    fixture for all 100 standard graphs of every size 6..16."""
    import json, os
    from quantumflow_amd.datasets import load_stdgraphs
    with open(os.path.join(os.path.dirname(__file__), "golden", "stdgraph_maxcut.json")) as f:
        want = json.load(f)["sizes"]
    graphs, best = [], []
    for n in range(6, 17):
        graphs += load_stdgraphs(n)
        best += want[str(n)]["maxcut"]
    print("maxcut batch of", len(graphs), "standard graphs", flush=True)
    batch = E.QAOABatch(graphs, 1, mode=E.MODE_FORWARD)
    assert np.array_equal(batch.maxcut, np.asarray(best, dtype=float))


def test_engine_values_on_standard_graphs(E):
    """QAOA values on standard 8- and 12-node graphs equal the gate-by-gate
    oracle (qubo_circuit + Circuit.run) to 1e-10."""
    from quantumflow_amd.datasets import load_stdgraphs
    graphs = load_stdgraphs(8)[:6] + load_stdgraphs(12)[:6]
    rng = np.random.default_rng(7)
    p = 2
    beta = rng.uniform(0, np.pi / 2, size=(len(graphs), p))
    gamma = rng.uniform(0, np.pi, size=(len(graphs), p))
    batch = E.QAOABatch(graphs, p, mode=E.MODE_FORWARD)
    val, _ = batch.evaluate(beta, gamma)
    batch.check()
    val = val.cpu().numpy()
    for i, g in enumerate(graphs):
        ref = O.qaoa_expectation(g.number_of_nodes(), list(g.edges()), beta[i], gamma[i])
        assert abs(val[i] - ref) <= 1e-10


def test_program_runs_with_another_block_size(E):
    """Forward programs move the RX bits a thread of their block size holds
    in registers into tile loads and stores.  Run with another block size,
    the kernel applies the bits it does not hold as LDS chunks, so results
    match the unmerged program."""
    import torch
    rng = np.random.default_rng(12)
    graphs = [er(n, 0.5, 120 + n) for n in (9, 14, 16)]
    p = 2
    beta = rng.uniform(0, np.pi / 2, size=(len(graphs), p))
    gamma = rng.uniform(0, np.pi, size=(len(graphs), p))
    ref, _ = E.QAOABatch(graphs, p, mode=E.MODE_FORWARD, tile_bits=13, block=512,
                         fuse=False).evaluate(beta, gamma)
    ref = ref.clone()
    for built, run in ((512, 256), (256, 512), (256, 1024)):
        print("program for block", built, "run with", run, flush=True)
        b = E.QAOABatch(graphs, p, mode=E.MODE_FORWARD, tile_bits=13, block=built)
        assert any(st[0] in (E.ST_RX_LOAD, E.ST_RX_STORE)
                   for _, sts, _ in E.build_program(16, p, E.MODE_FORWARD, 13, True, built)
                   for st in sts)
        b._args.block = run
        v, _ = b.evaluate(beta, gamma)
        b.check()
        assert torch.max(torch.abs(v - ref)).item() <= 1e-12


@pytest.mark.parametrize("mode", [0, 1])
def test_mirror_half_state_matches_full_state(E, mode):
    """Instances run on half their state (complement symmetry of MaxCut:
    psi(z) = psi(~z), items only over tiles whose highest non-tile bit is
    clear, results counted twice) give the values and gradients of the
    whole-state program, weighted graphs included; the mirror pools hold
    half the amplitudes."""
    rng = np.random.default_rng(12 + mode)
    graphs = [er(n, 0.45, 200 + n) for n in (15, 16, 18, 21)]
    g = er(17, 0.5, 77)
    for u, v in g.edges():
        g[u][v]['weight'] = float(rng.uniform(0.1, 2.0))
    graphs.append(g)
    p = 3
    beta = rng.uniform(0, np.pi / 2, size=(len(graphs), p))
    gamma = rng.uniform(0, np.pi, size=(len(graphs), p))
    half = E.QAOABatch(graphs, p, mode=mode)
    full = E.QAOABatch(graphs, p, mode=mode, mirror=False)
    assert half.psi_pool.numel() < full.psi_pool.numel()
    v1, g1 = half.evaluate(beta, gamma)
    v2, g2 = full.evaluate(beta, gamma)
    half.check()
    full.check()
    assert np.max(np.abs(v1.cpu().numpy() - v2.cpu().numpy())) <= 1e-11
    if mode == 1:
        assert np.max(np.abs(g1.cpu().numpy() - g2.cpu().numpy())) <= 1e-11
    # and the oracle, for the largest unweighted instance
    i = 3
    ref = O.qaoa_expectation_fast(len(graphs[i]), list(graphs[i].edges()), beta[i], gamma[i])
    assert abs(v1.cpu().numpy()[i] - ref) <= 1e-10

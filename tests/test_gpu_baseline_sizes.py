"""Parity at the benchmark sizes themselves.

* 28 qubits, depth 40, seed 2024 (the headline circuit, 4 GiB state): the
  fused plan the benchmark runs against the unfused per-gate path, both from
  |0...0>: max |d amp| <= 1e-12 and |<psi|psi> - 1| <= 1e-12.  The unfused
  path is itself checked against the oracle gate by gate at 20 qubits.
* 512 ER graphs, n 10..22, p = 4 (the batched values benchmark): the whole
  batch in one launch; six n = 20..22 instances and every n <= 12 instance
  against the oracle, <= 1e-10.
* 64 3-regular graphs, n 16..24, p = 8 (the value + gradient benchmark):
  one n = 24 instance's value and 16 gradient entries against a torch
  complex128 autograd restatement run on the GPU (<= 1e-10), its value
  against the numpy oracle (<= 1e-10), and two entries against the
  4th-order central difference at h = 1e-3 (<= 1e-9).
"""
import numpy as np
import pytest

from oracle import qf_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qf():
    import quantumflow_amd as qf
    assert qf.backend.gpu_available()
    return qf


def test_headline_28q_fused_vs_unfused(qf):
    import torch
    from quantumflow_amd import fusion
    from quantumflow_amd.workloads import random_mixed_circuit, build_gates
    n = 28
    qubits = tuple(range(n))
    gates = build_gates(random_mixed_circuit(n, 40, seed=2024))
    plan = fusion.plan_for(gates, qubits)
    print("28q headline: %d gates, %d passes, mode %s" % (len(gates), plan.n_passes, plan.mode),
          flush=True)
    work = torch.zeros(2 ** n, dtype=torch.complex128, device="cuda")
    work[0] = 1.0
    plan.run_inplace(work)
    ref = torch.zeros_like(work)
    ref[0] = 1.0
    fusion.apply_gates_inplace(gates, qubits, ref)
    torch.cuda.synchronize()
    err = float((work - ref).abs().max().item())
    norm = float(qf.backend.norm(work))
    print("28q headline: max |d amp| = %.3e, |norm - 1| = %.3e" % (err, abs(norm - 1)),
          flush=True)
    assert err <= 1e-12
    assert abs(norm - 1.0) <= 1e-12
    del work, ref
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def config3():
    from quantumflow_amd import qaoa_engine as E
    from quantumflow_amd.workloads import qaoa_batch_instances
    inst = qaoa_batch_instances(512, 10, 22, 4, seed=11)
    graphs = [g for g, _, _ in inst]
    beta = np.array([b for _, b, _ in inst])
    gamma = np.array([g for _, _, g in inst])
    batch = E.QAOABatch(graphs, 4, mode=E.MODE_FORWARD)
    batch.set_angles(beta, gamma)
    batch.capture()
    batch.replay()
    batch.check()
    vals = batch.out.view(-1).cpu().numpy()
    return graphs, beta, gamma, vals


def _large_config3(graphs):
    big = [i for i, g in enumerate(graphs) if len(g) >= 20]
    # two of each size 20, 21, 22
    pick = []
    for n in (20, 21, 22):
        pick += [i for i in big if len(graphs[i]) == n][:2]
    return pick


@pytest.mark.parametrize("slot", range(6))
def test_config3_large_instance_vs_oracle(qf, config3, slot):
    graphs, beta, gamma, vals = config3
    i = _large_config3(graphs)[slot]
    g = graphs[i]
    n = len(g)
    ref = O.qaoa_expectation_fast(n, list(g.edges()), beta[i], gamma[i])
    print("config 3 instance %d: n=%d |E|=%d engine %.15f oracle %.15f"
          % (i, n, g.number_of_edges(), vals[i], ref), flush=True)
    assert abs(vals[i] - ref) <= 1e-10


def test_config3_small_instances_vs_oracle(qf, config3):
    graphs, beta, gamma, vals = config3
    small = [i for i, g in enumerate(graphs) if len(g) <= 12]
    assert len(small) > 50
    worst = 0.0
    for i in small:
        g = graphs[i]
        ref = O.qaoa_expectation_fast(len(g), list(g.edges()), beta[i], gamma[i])
        worst = max(worst, abs(vals[i] - ref))
    print("config 3: %d instances with n <= 12, max |d<C>| = %.3e" % (len(small), worst),
          flush=True)
    assert worst <= 1e-10


def test_config4_value_grad_n24(qf):
    import torch
    from quantumflow_amd import qaoa_engine as E
    from quantumflow_amd.workloads import regular_instances
    inst = regular_instances(64, 16, 24, 8, seed=12)
    graphs = [g for g, _, _ in inst]
    beta = np.array([b for _, b, _ in inst])
    gamma = np.array([g for _, _, g in inst])
    i = next(k for k, g in enumerate(graphs) if len(g) == 24)
    g = graphs[i]
    n, edges, p = 24, list(g.edges()), 8
    batch = E.QAOABatch(graphs, p, mode=E.MODE_BACKWARD)
    val, grad = batch.evaluate(beta, gamma)
    batch.check()
    val = float(val[i].item())
    grad = grad[i].cpu().numpy()
    del batch
    torch.cuda.empty_cache()
    print("config 4 instance %d: n=24, engine value %.15f" % (i, val), flush=True)

    b = torch.tensor(beta[i], dtype=torch.float64, device="cuda", requires_grad=True)
    gm = torch.tensor(gamma[i], dtype=torch.float64, device="cuda", requires_grad=True)
    v = O.qaoa_value_torch(n, edges, b, gm, device="cuda")
    v.backward()
    tg = np.concatenate([b.grad.cpu().numpy(), gm.grad.cpu().numpy()])
    print("config 4: torch autograd value %.15f, max |d grad| = %.3e"
          % (v.item(), np.max(np.abs(grad - tg))), flush=True)
    assert abs(val - v.item()) <= 1e-10
    assert np.max(np.abs(grad - tg)) <= 1e-10
    del v, b, gm
    torch.cuda.empty_cache()

    # 4th-order central difference on two angles (beta_0, gamma_7), with the
    # torch restatement as the function
    def f(bb, gg):
        with torch.no_grad():
            return O.qaoa_value_torch(n, edges, torch.tensor(bb, device="cuda"),
                                      torch.tensor(gg, device="cuda"), device="cuda").item()
    h = 1e-3
    for which, j in (("beta", 0), ("gamma", 7)):
        def ev(d):
            bb, gg = beta[i].copy(), gamma[i].copy()
            (bb if which == "beta" else gg)[j] += d
            return f(bb, gg)
        fd = (-ev(2 * h) + 8 * ev(h) - 8 * ev(-h) + ev(-2 * h)) / (12 * h)
        k = j if which == "beta" else p + j
        print("config 4: d/d%s_%d engine %.12f fd4 %.12f" % (which, j, grad[k], fd), flush=True)
        assert abs(grad[k] - fd) <= 1e-9
    torch.cuda.empty_cache()

    ref = O.qaoa_expectation_fast(n, edges, beta[i], gamma[i])
    print("config 4: numpy oracle value %.15f" % ref, flush=True)
    assert abs(val - ref) <= 1e-10

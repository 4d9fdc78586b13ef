#!/usr/bin/env python
"""Benchmark: 28-qubit complex128 random-circuit throughput on one MI355X.

Workload: the seeded 28-qubit, depth-40 random circuit (RX/RY/RZ/PHASE 50%,
CNOT/CZ/ISWAP/SWAP 45%, CCNOT 5%, targets uniform over all qubits), state
resident in HBM (4 GiB).  One step = the whole circuit applied once, in
place, to the working state.  Prints one JSON line:

  value      gates/s of the default executor (fused when available);
             gates_per_s_pct_of_hbm_roofline prices every gate at one state
             read + write (the unfused cost), passes_per_s_pct_of_hbm_roofline
             every fused pass
  roofline   dominant kernel: algorithmic bytes per launch (32 * 2^28 B: one
             read + one write of the state) / its average duration, timed
             with HIP events on the launching stream inside the timed region
  fused_mfma_blocks  the circuit fused into dense 3/4/5-qubit blocks on the
             matrix cores (FP64 MFMA)
  cpu_baseline  the numpy oracle (np.einsum per gate, 1 thread) on a
             bounded prefix of the same circuit, on this host

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

N_QUBITS = 28
DEPTH = 40
SEED = 2024
HBM_PEAK_GBS = 8000.0
METRIC = "28q c128 gates/s as % HBM roofline; QAOA value+grad/s vs plain, same GPU"


def cpu_baseline(desc, n, ngates):
    """Time the oracle (numpy einsum per gate) on the first ``ngates`` gates
    of the circuit at the full size, on this host."""
    import numpy as np
    from oracle import qf_oracle as O
    ket = np.zeros(2 ** n, dtype=np.complex128)
    ket[0] = 1.0
    ket = ket.reshape([2] * n)
    qubits = tuple(range(n))
    t0 = time.perf_counter()
    for g in desc[:ngates]:
        ket = O.run_gate(ket, qubits, O.gate_matrix(g[0], *g[1]), g[2])
    dt = time.perf_counter() - t0
    del ket
    return {"value": ngates / dt, "unit": "gates/s", "cores": 1, "kind": "port",
            "sample": "first %d gates of the %d-qubit depth-%d circuit, numpy einsum "
                      "per gate (single-threaded), host os.cpu_count()=%d, affinity=%d"
                      % (ngates, n, DEPTH, os.cpu_count(), len(os.sched_getaffinity(0)))}


def measured_traffic(plan, n):
    """HBM bytes per tile-pass launch from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE runs (scaled by a plain copy of known size), when
    they were taken on this workload and plan; else None."""
    path = os.path.join(HERE, "profiles", "tile_pass_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if plan is None or (n, DEPTH, SEED) != (rec.get("qubits"), rec.get("depth"), rec.get("seed")):
        return None
    st = plan.stats()
    if any(st.get(k) != v for k, v in rec.get("plan", {}).items()):
        return None
    return rec.get("traffic_bytes_per_launch")


def fused_block_section(gates, qubits, work, widths=(3, 4, 5), reps=2):
    """The same circuit with every pass fused into dense 3..K-qubit blocks
    applied on the matrix cores (FP64 MFMA): gates/s, blocks/s, the tile
    pass's HBM rate and the MFMA rate (real-form flops: 8 * 2^k per
    amplitude and block)."""
    import torch
    from quantumflow_amd import fusion
    n = len(qubits)
    out = {}
    for k in widths:
        plan = fusion.TilePlan(gates, qubits, tile_bits=12, low_bits=3, block=256,
                               fuse_width=k, mode="mfma", persistent=True)
        if plan.mode != "mfma":
            continue
        plan.run_inplace(work)
        prof = fusion.KernelTimer()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.run_inplace(work, prof=prof)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        name, avg_ms, _ = prof.dominant()
        flops = sum(8.0 * 2 ** (n + o.k) for o in plan.ops_host if o.kind == fusion.OP_MFMA)
        out["k%d" % k] = {
            "gates_per_s": round(len(gates) / dt, 2), "ms_per_circuit": round(dt * 1e3, 3),
            "blocks": plan.n_blocks, "passes": plan.n_passes,
            "blocks_per_s": round(plan.n_blocks / dt, 1),
            "tile_pass_avg_ms": round(avg_ms, 4),
            "tile_pass_GBs": round(32.0 * 2 ** n / (avg_ms * 1e-3) / 1e9, 1),
            "mfma_TFLOPs": round(flops / dt / 1e12, 2), "fp64_peak_TFLOPs": 78.6}
    return out


def density_section(n=13, ngates=40, seed=5):
    """Density-matrix evolution: an n-qubit density (a 2n-axis tensor,
    16 * 4^n bytes) through a random circuit of 1- and 2-qubit gates, then a
    depolarizing layer on every qubit.  A gate is one gate-kernel pass over
    the tensor (U (x) conj U on the row and column axes) or two when no
    kernel takes that product; a 4-operator channel is four gate evolutions
    plus four weighted sums.  Synthetic data; HIP events around the whole
    sequence."""
    from quantumflow_amd.channels import evolve_passes
    import numpy as np
    import torch
    import quantumflow_amd as qf

    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(ngates):
        if rng.random() < 0.5:
            gates.append(qf.RX(float(rng.uniform(0, 6)), int(rng.integers(n))))
        else:
            a, b = rng.choice(n, size=2, replace=False)
            gates.append(qf.CNOT(int(a), int(b)))
    rho = qf.mixed_density(n)
    rho = qf.RY(0.3, 0).evolve(rho)                  # warm the kernels
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for g in gates:
        rho = g.evolve(rho)
    gate_end = torch.cuda.Event(enable_timing=True)
    gate_end.record()
    for q in range(n):
        rho = qf.Depolarizing(0.01, q).evolve(rho)
    end.record()
    torch.cuda.synchronize()
    gate_ms = start.elapsed_time(gate_end)
    chan_ms = gate_end.elapsed_time(end)
    tr = complex(rho.trace().item())
    nbytes = 16.0 * 4 ** n
    return {"workload": "%d-qubit density (%.2f GiB), %d random RX/CNOT gates, then "
                        "depolarizing on every qubit" % (n, nbytes / 2 ** 30, ngates),
            "gate_evolutions_per_s": round(ngates / (gate_ms * 1e-3), 1),
            "ms_per_gate_evolution": round(gate_ms / ngates, 4),
            "passes_per_gate": round(sum(evolve_passes(g) for g in gates) / ngates, 3),
            "gate_pass_GBs": round(2 * nbytes * sum(evolve_passes(g) for g in gates)
                                   / (gate_ms * 1e-3) / 1e9, 1),
            "ms_per_channel": round(chan_ms / n, 4),
            "trace_after": round(tr.real, 12)}


def qaoa_section(steps, plain_sample=24, grad_sample=4):
    """Batched QAOA (512 ER graphs, n 10..22, p=4: values; 64 3-regular
    graphs, n 16..24, p=8: values + adjoint gradients) against the plain
    per-gate GPU path on samples of the same instances."""
    import numpy as np
    import torch
    from quantumflow_amd import qaoa as Q
    from quantumflow_amd import qaoa_engine as E
    from quantumflow_amd.workloads import qaoa_batch_instances, regular_instances

    def timed(fn, reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    out = {}
    # values: 512 ER instances
    inst = qaoa_batch_instances(512, 10, 22, 4, seed=11)
    graphs = [g for g, _, _ in inst]
    beta = np.array([b for _, b, _ in inst])
    gamma = np.array([g for _, _, g in inst])
    t0 = time.perf_counter()
    batch = E.QAOABatch(graphs, 4, mode=E.MODE_FORWARD)
    batch.set_angles(beta, gamma)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    batch.capture()
    dt = timed(batch.replay, steps)
    batch.check()
    vals = batch.out.view(-1).cpu().numpy()
    tp = timed(lambda: [Q.plain_value(graphs[i], beta[i], gamma[i]) for i in range(plain_sample)], 1)
    plain_vals = [Q.plain_value(graphs[i], beta[i], gamma[i]).item() for i in range(plain_sample)]
    out["forward"] = {
        "workload": "512 ER graphs n=10..22 p_edge=0.3..0.6, p=4, complex128",
        "batched_instances_per_s": round(512 / dt, 1), "batched_ms": round(dt * 1e3, 3),
        "setup_s (programs + cut tables, once per graph set)": round(setup, 3),
        "plain_instances_per_s": round(plain_sample / tp, 2),
        "plain_sample": plain_sample,
        "speedup_vs_plain": round((512 / dt) / (plain_sample / tp), 1),
        "max_abs_diff_vs_plain": float(np.max(np.abs(vals[:plain_sample] - np.array(plain_vals)))),
    }
    # values + gradients: 64 3-regular instances, p=8
    inst = regular_instances(64, 16, 24, 8, seed=12)
    graphs = [g for g, _, _ in inst]
    beta = np.array([b for _, b, _ in inst])
    gamma = np.array([g for _, _, g in inst])
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    batch = E.QAOABatch(graphs, 8, mode=E.MODE_BACKWARD)
    batch.set_angles(beta, gamma)
    batch.capture()
    dt = timed(batch.replay, steps)
    batch.check()
    peak = torch.cuda.max_memory_allocated() - base
    res = batch.out.view(64, -1).cpu().numpy()
    errs, tps = [], 0.0
    for i in range(grad_sample):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        v, g = Q.plain_value_and_grad(graphs[i], beta[i], gamma[i])
        torch.cuda.synchronize()
        tps += time.perf_counter() - t0
        errs.append(float(np.max(np.abs(g.cpu().numpy() - res[i, 1:]))))
        errs.append(abs(v.item() - res[i, 0]))
    stored = sum((len(g) + 8 * (g.number_of_edges() + len(g))) * 16 * 2 ** len(g) for g in graphs)
    out["value_and_grad"] = {
        "workload": "64 random 3-regular graphs n=16..24 (even), p=8, complex128",
        "batched_value_grad_per_s": round(64 / dt, 2), "batched_ms": round(dt * 1e3, 3),
        "plain_value_grad_per_s": round(grad_sample / tps, 3), "plain_sample": grad_sample,
        "speedup_vs_plain": round((64 / dt) / (grad_sample / tps), 1),
        "max_abs_diff_vs_plain_adjoint": max(errs),
        "engine_peak_bytes": int(peak),
        "stored_intermediates_bytes (one state per gate, as autograd keeps)": int(stored),
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=N_QUBITS)
    ap.add_argument("--executor", choices=["auto", "fused", "unfused"], default="auto")
    ap.add_argument("--cpu-gates", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-qaoa", action="store_true")
    ap.add_argument("--no-density", action="store_true")
    ap.add_argument("--no-mfma", action="store_true")
    ap.add_argument("--no-unfused", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain device copy of the state (4 GiB read + 4 GiB "
                         "written at 28 qubits) to scale HBM byte counters against")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import quantumflow_amd as qf
    from quantumflow_amd import fusion
    from quantumflow_amd.workloads import random_mixed_circuit, build_gates

    n = args.qubits
    desc = random_mixed_circuit(n, DEPTH, seed=SEED)
    gates = build_gates(desc)
    qubits = tuple(range(n))
    print("# workload: %d gates on %d qubits" % (len(gates), n), file=sys.stderr, flush=True)

    ket = qf.zero_state(n)
    work = qf.backend.as_device(ket.tensor).clone()
    state = qf.State(work.view([2] * n), qubits)

    plan = None
    if args.executor in ("auto", "fused"):
        plan = fusion.plan_for(gates, qubits)
        if plan is None and args.executor == "fused":
            raise SystemExit("fused executor unavailable")

    def step(prof=None):
        if plan is not None:
            plan.run_inplace(work, prof=prof)
        else:
            fusion.apply_gates_inplace(gates, qubits, work, prof)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    prof = fusion.KernelTimer()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(prof)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    gates_per_step = len(gates)
    value = gates_per_step * args.steps / dt
    ms = dt / args.steps * 1e3
    bytes_per_pass = 32.0 * 2 ** n
    kname, avg_ms, launches = prof.dominant()
    achieved = bytes_per_pass / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(plan, n),
            "kernel": kname, "avg_ms": round(avg_ms, 4), "launches_timed": launches,
            "bytes_per_launch": bytes_per_pass}

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "gates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex128", "data": "synthetic",
        "config": {"workload": "random_mixed_circuit", "qubits": n, "depth": DEPTH,
                   "seed": SEED, "gates": gates_per_step,
                   "executor": "fused" if plan is not None else "unfused",
                   "passes_per_step": (plan.n_passes if plan is not None else gates_per_step)},
        "gates_per_s_pct_of_hbm_roofline": round(100.0 * value * bytes_per_pass / (HBM_PEAK_GBS * 1e9), 2),
        "passes_per_s_pct_of_hbm_roofline": round(
            100.0 * value / gates_per_step * (plan.n_passes if plan is not None else gates_per_step)
            * bytes_per_pass / (HBM_PEAK_GBS * 1e9), 2),
        "roofline": roof,
    }
    if plan is not None:
        out["config"].update(plan.stats())
    if not args.no_unfused:
        unf = fusion.KernelTimer()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fusion.apply_gates_inplace(gates, qubits, work, unf)
        torch.cuda.synchronize()
        dtu = time.perf_counter() - t0
        un_name, un_ms, _ = unf.dominant()
        out["unfused"] = {"gates_per_s": round(len(gates) / dtu, 2),
                          "dominant_kernel": un_name, "avg_ms": round(un_ms, 4),
                          "achieved_GBs": round(bytes_per_pass / (un_ms * 1e-3) / 1e9, 1)}
    if args.copy_calibration:
        spare = torch.empty_like(work)
        spare.copy_(work)
        torch.cuda.synchronize()
        del spare
    if not args.no_mfma:
        out["fused_mfma_blocks"] = fused_block_section(gates, qubits, work)
    if not args.no_qaoa:
        del state
        out["qaoa"] = qaoa_section(max(2, args.steps))
    if not args.no_density:
        out["density"] = density_section()
    if not args.no_cpu and args.cpu_gates > 0:
        del work
        torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline(desc, n, args.cpu_gates)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

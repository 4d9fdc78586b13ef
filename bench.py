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
  tile_pass_fp64  the same launches against the FP64 peak: the plan's
             algorithmic floating-point operations per launch / duration
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
FP64_PEAK_TFLOPS = 78.6
METRIC = "28q c128 gates/s as % HBM roofline; QAOA value+grad/s vs plain, same GPU"


def note(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print("# " + msg, file=sys.stderr, flush=True)


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


def qaoa_section(steps, grad_compare_n=18):
    """Batched QAOA (512 ER graphs, n 10..22, p=4: values; 64 3-regular
    graphs, n 16..24, p=8: values + adjoint gradients) against the plain
    per-gate GPU path run on exactly the same instances, one at a time; and
    one n=18, p=8 instance's value + gradient by stored-intermediates torch
    autograd on the GPU against the engine (peak memory and time)."""
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
    del batch
    plain_vals = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(len(graphs)):
        plain_vals.append(Q.plain_value(graphs[i], beta[i], gamma[i]))
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    plain_vals = torch.stack(plain_vals).cpu().numpy()
    out["forward"] = {
        "workload": "512 ER graphs n=10..22 p_edge=0.3..0.6, p=4, complex128",
        "batched_instances_per_s": round(512 / dt, 1), "batched_ms": round(dt * 1e3, 3),
        "setup_s (programs + cut tables, once per graph set)": round(setup, 3),
        "plain_instances_per_s": round(len(graphs) / tp, 2), "plain_s": round(tp, 3),
        "plain_sample": "all 512 instances, one at a time (same instances as the batch)",
        "speedup_vs_plain": round((512 / dt) / (len(graphs) / tp), 1),
        "max_abs_diff_vs_plain": float(np.max(np.abs(vals - plain_vals))),
    }
    # values + gradients: 64 3-regular instances, p=8
    inst = regular_instances(64, 16, 24, 8, seed=12)
    graphs = [g for g, _, _ in inst]
    beta = np.array([b for _, b, _ in inst])
    gamma = np.array([g for _, _, g in inst])
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    batch = E.QAOABatch(graphs, 8, mode=E.MODE_BACKWARD)
    batch.set_angles(beta, gamma)
    batch.capture()
    dt = timed(batch.replay, steps)
    batch.check()
    peak = torch.cuda.max_memory_allocated() - base
    res = batch.out.view(64, -1).cpu().numpy()
    del batch
    errs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(len(graphs)):
        v, g = Q.plain_value_and_grad(graphs[i], beta[i], gamma[i])
        errs.append(float(np.max(np.abs(g.cpu().numpy() - res[i, 1:]))))
        errs.append(abs(v.item() - res[i, 0]))
    torch.cuda.synchronize()
    tps = time.perf_counter() - t0
    stored = sum((len(g) + 8 * (g.number_of_edges() + len(g))) * 16 * 2 ** len(g) for g in graphs)
    out["value_and_grad"] = {
        "workload": "64 random 3-regular graphs n=16..24 (even), p=8, complex128",
        "batched_value_grad_per_s": round(64 / dt, 2), "batched_ms": round(dt * 1e3, 3),
        "plain_value_grad_per_s": round(len(graphs) / tps, 3), "plain_s": round(tps, 3),
        "plain_sample": "all 64 instances, one at a time: per-gate forward, per-gate adjoint sweep",
        "speedup_vs_plain": round((64 / dt) / (len(graphs) / tps), 1),
        "max_abs_diff_vs_plain_adjoint": max(errs),
        "engine_peak_bytes": int(peak),
        "stored_intermediates_bytes_estimate (one state per gate)": int(stored),
    }
    # one instance: stored-intermediates autograd vs the engine, measured
    k = next(i for i, g in enumerate(graphs) if len(g) == grad_compare_n)
    g1 = [graphs[k]]
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    t0 = time.perf_counter()
    av, ag = Q.autograd_value_and_grad(graphs[k], beta[k], gamma[k])
    torch.cuda.synchronize()
    t_auto = time.perf_counter() - t0
    peak_auto = torch.cuda.max_memory_allocated() - base
    av, ag = av.item(), ag.cpu().numpy()
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    one = E.QAOABatch(g1, 8, mode=E.MODE_BACKWARD)
    ev, eg = one.evaluate(beta[k:k + 1], gamma[k:k + 1])
    one.check()
    t_eng = timed(lambda: one.evaluate(beta[k:k + 1], gamma[k:k + 1]), 5)
    peak_eng = torch.cuda.max_memory_allocated() - base
    ev, eg = ev[0].item(), eg[0].cpu().numpy()
    del one
    out["stored_intermediates_backward"] = {
        "workload": "one random 3-regular graph n=%d (|E|=%d), p=8: value + 16-entry gradient"
                    % (grad_compare_n, graphs[k].number_of_edges()),
        "autograd_peak_bytes": int(peak_auto), "autograd_s": round(t_auto, 4),
        "engine_peak_bytes": int(peak_eng), "engine_s": round(t_eng, 5),
        "memory_ratio": round(peak_auto / max(peak_eng, 1), 1),
        "max_abs_diff_grad": float(np.max(np.abs(ag - eg))),
        "abs_diff_value": abs(av - ev),
    }
    return out


def qaoa_batches_only(steps, pass_times=False):
    """Only the two batched QAOA launches (graph replays), for profilers:
    no plain-path comparison, no other section.  With ``pass_times``, one
    more launch of each with the kernel's wall-clock stamps on: per global
    pass, the span and the workgroups' mean and largest busy time; what the
    spans hold beyond the mean busy time is barrier wait and imbalance."""
    import numpy as np
    import torch
    from quantumflow_amd import qaoa_engine as E
    from quantumflow_amd.workloads import qaoa_batch_instances, regular_instances
    out = {}
    for name, inst, p, mode in (
            ("forward", qaoa_batch_instances(512, 10, 22, 4, seed=11), 4, E.MODE_FORWARD),
            ("value_and_grad", regular_instances(64, 16, 24, 8, seed=12), 8, E.MODE_BACKWARD)):
        graphs = [g for g, _, _ in inst]
        batch = E.QAOABatch(graphs, p, mode=mode)
        batch.set_angles(np.array([b for _, b, _ in inst]), np.array([g for _, _, g in inst]))
        batch.capture()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.replay()
        torch.cuda.synchronize()
        batch.check()
        ms = (time.perf_counter() - t0) / steps * 1e3
        sb, cb = batch.algorithmic_bytes()
        out[name] = {"instances": len(graphs), "ms_per_batch": round(ms, 3),
                     "algorithmic_state_GB": round(sb / 1e9, 3),
                     "algorithmic_cut_GB": round(cb / 1e9, 3),
                     "algorithmic_TBs": round((sb + cb) / (ms * 1e-3) / 1e12, 3)}
        if pass_times:
            pt = batch.pass_times()
            span = sum(q["span_us"] for q in pt["passes"])
            busy = sum(q["busy_mean_us"] for q in pt["passes"])
            out[name]["pass_times"] = {
                "grid": pt["grid"], "global_passes": len(pt["passes"]),
                "total_us": pt["total_us"], "sum_span_us": round(span, 1),
                "sum_busy_mean_us": round(busy, 1),
                "sum_busy_max_us": round(sum(q["busy_max_us"] for q in pt["passes"]), 1),
                "barrier_and_imbalance_frac": round((span - busy) / span, 4),
                "passes": pt["passes"]}
        del batch
        torch.cuda.empty_cache()
    return out


def subop_costs(n=N_QUBITS, reps=5):
    """Time per register-phase sub-op kind: one-pass plans on an n-qubit
    state whose gates all act on the tile's bits (index bits 3..11 as
    targets; controls and phase qubits on bits 20..27, outside the tile),
    16 and 48 gates of one kind, HIP events around `reps` runs.  The
    difference between the two sizes over the difference in sub-ops is the
    cost of one more sub-op of that kind."""
    import numpy as np
    import torch
    import quantumflow_amd as qf
    from quantumflow_amd import backend as bk
    from quantumflow_amd import fusion
    from collections import Counter
    rng = np.random.default_rng(0)

    def q(b):                       # index bit -> qubit
        return n - 1 - b
    names = {fusion.SUB_D1: "D1", fusion.SUB_D1R: "D1R", fusion.SUB_D1X: "D1X",
             fusion.SUB_DIAG: "DIAG", fusion.SUB_DIAGV: "DIAGV", fusion.SUB_CX: "CX"}

    def gates(kind, k):
        out = []
        for i in range(k):
            t = q(3 + i % 9)
            if kind == "rx":
                out.append(qf.RX(float(rng.uniform(0, 6)), t))
            elif kind == "ry":
                out.append(qf.RY(float(rng.uniform(0, 6)), t))
            elif kind == "cx_in_registers":
                a, b = rng.choice(9, 2, replace=False) + 3
                out.append(qf.CNOT(q(int(a)), q(int(b))))
            elif kind == "cx_control_outside_tile":
                out.append(qf.CNOT(q(20 + i % 8), t))
            elif kind == "cz_outside_tile":
                out.append(qf.CZ(q(20 + i % 8), t))
            elif kind == "rx_carrying_a_flip":
                # a CNOT controlled outside the tile, then an RX on another
                # bit: the CNOT becomes a lane flip riding on the RX record
                out.append(qf.CNOT(q(20 + i % 8), t))
                out.append(qf.RX(float(rng.uniform(0, 6)), q(3 + (i + 4) % 9)))
        return out
    work = bk.zero_state_tensor(n).reshape(-1).clone()
    out = {}
    # (lone CNOTs controlled outside the tile have no record to ride on as
    # lane flips: they stay exchange records)
    for kind in ("rx", "ry", "cx_in_registers", "cx_control_outside_tile", "cz_outside_tile",
                 "rx_carrying_a_flip"):
        row = {}
        for k in (16, 48):
            # records kept apart (no merging), so that each gate is one record
            plan = fusion.TilePlan(gates(kind, k), tuple(range(n)), tile_bits=12, low_bits=3,
                                   block=256, fuse_width=1, mode="reg", merge_records=False)
            subs = Counter(names[plan.srcs[o.src + 4 * j] & fusion.REC_TYPE] for o in plan.ops_host
                           if o.kind == fusion.OP_RPHASE for j in range(o.coef))
            flips = sum(bool(plan.srcs[o.src + 4 * j] >> 8) for o in plan.ops_host
                        if o.kind == fusion.OP_RPHASE for j in range(o.coef))
            if flips:
                subs["lane flips riding"] = flips
            plan.run_inplace(work)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                plan.run_inplace(work)
            e1.record()
            torch.cuda.synchronize()
            row["gates_%d" % k] = {"ms_per_pass": round(e0.elapsed_time(e1) / reps / plan.n_passes, 4),
                                   "passes": plan.n_passes, "phases": plan.n_phases,
                                   "sub_ops": dict(subs)}
        a, b = row["gates_16"], row["gates_48"]
        extra = sum(v for k2, v in b["sub_ops"].items() if k2 in names.values()) - \
            sum(v for k2, v in a["sub_ops"].items() if k2 in names.values())
        if extra > 0 and a["passes"] == b["passes"] == 1:
            row["us_per_extra_sub_op"] = round((b["ms_per_pass"] - a["ms_per_pass"]) * 1e3 / extra, 1)
        out[kind] = row
    return out


def config1_section(reps=20):
    """The 12-qubit, 20-layer rotation circuit (360 gates) from |0...0>:
    the state is 64 KiB, so this is launch- and latency-bound: microseconds
    per circuit and gates/s, on the per-gate path and the fused executor
    (one or two passes), plus all <Z_i>."""
    import torch
    import quantumflow_amd as qf
    from quantumflow_amd import fusion
    from quantumflow_amd.workloads import layered_rotation_circuit, build_gates
    n = 12
    gates = build_gates(layered_rotation_circuit(n, 20, seed=1))
    qubits = tuple(range(n))
    circ = qf.Circuit(gates)
    out = {"workload": "12-qubit layered rotation circuit, 20 layers, %d gates, seed 1" % len(gates)}

    def per_circuit(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    dt = per_circuit(lambda: circ.run(qf.zero_state(n), fused=False).expectation_z())
    out["per_gate_us_per_circuit"] = round(dt * 1e6, 1)
    out["per_gate_gates_per_s"] = round(len(gates) / dt, 1)
    plan = fusion.plan_for(gates, qubits)
    work = qf.backend.as_device(qf.zero_state(n).tensor).reshape(-1).clone()

    def fused():
        work.zero_()
        work[0] = 1.0
        plan.run_inplace(work)
        qf.State(work.view([2] * n), qubits).expectation_z()
    dt = per_circuit(fused)
    out["fused_us_per_circuit"] = round(dt * 1e6, 1)
    out["fused_gates_per_s"] = round(len(gates) / dt, 1)
    out["fused_passes"] = plan.n_passes
    return out


def gops_section(sizes=(16, 20, 24, 28), gates=100, reps=3):
    """The reference's GOPS harness (tools/benchmark.py:31-53): H on every
    qubit, then X, T, CNOT triples on random pairs, GOPS = 100 / seconds.
    Gate by gate through Gate.run (the reference's call pattern: each gate
    returns a new state) and through Circuit.run's fused executor."""
    import torch
    import quantumflow_amd as qf
    from quantumflow_amd.workloads import gops_circuit, build_gates
    out = {}
    for n in sizes:
        gl = build_gates(gops_circuit(n, gates, seed=n))
        circ = qf.Circuit(gl)

        def per_gate():
            ket = qf.zero_state(n)
            for g in gl:
                ket = g.run(ket)
            return ket

        def fused():
            return circ.run(qf.zero_state(n))
        row = {}
        for name, fn in (("per_gate", per_gate), ("fused", fused)):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / reps
            row[name + "_gops"] = round(gates / dt, 1)
        out["n%d" % n] = row
        torch.cuda.empty_cache()
    return out


def _qaoa_cpu_value(args):
    """One config-3 instance's <C> by the numpy oracle (one process)."""
    from oracle import qf_oracle as O
    n, edges, beta, gamma = args
    t0 = time.perf_counter()
    O.qaoa_expectation_fast(n, edges, beta, gamma)
    return n, time.perf_counter() - t0


def cpu_extra_baselines(desc20, desc24):
    """CPU legs beyond the headline sample, on this host with the oracle:
    config 1 (12 qubits, whole circuit, 1 thread); config 2 at 20 qubits
    (whole circuit) and 24 qubits (first 48 gates), numpy einsum per gate,
    1 thread; config 3 values on all cores (a process per instance) over 4
    instances per qubit count, extrapolated to the 512-instance mix by the
    per-n mean times; config 4 value + gradient by torch complex128 autograd
    on the CPU (intra-op threads = cores) for two n=16 instances."""
    import multiprocessing as mp
    import numpy as np
    import torch
    from oracle import qf_oracle as O
    from quantumflow_amd.workloads import (layered_rotation_circuit, qaoa_batch_instances,
                                           regular_instances)
    cores = len(os.sched_getaffinity(0))
    procs = min(16, cores)
    out = {"host_cpu_count": os.cpu_count(), "affinity": cores}

    def run(desc, n, ngates):
        ket = np.zeros(2 ** n, dtype=np.complex128)
        ket[0] = 1.0
        ket = ket.reshape([2] * n)
        qubits = tuple(range(n))
        t0 = time.perf_counter()
        for g in desc[:ngates]:
            ket = O.run_gate(ket, qubits, O.gate_matrix(g[0], *g[1]), g[2])
        return time.perf_counter() - t0

    d1 = layered_rotation_circuit(12, 20, seed=1)
    t = run(d1, 12, len(d1))
    out["config1_12q"] = {"us_per_circuit": round(t * 1e6, 1),
                          "gates_per_s": round(len(d1) / t, 1), "threads": 1}
    t = run(desc20, 20, len(desc20))
    out["config2_20q_full_circuit"] = {"s": round(t, 3), "gates": len(desc20),
                                       "gates_per_s": round(len(desc20) / t, 2), "threads": 1}
    t = run(desc24, 24, 48)
    out["config2_24q_first48"] = {"s": round(t, 3), "gates": 48,
                                  "gates_per_s": round(48 / t, 3), "threads": 1}
    inst = qaoa_batch_instances(512, 10, 22, 4, seed=11)
    by_n = {}
    for g, b, gm in inst:
        by_n.setdefault(len(g), []).append((len(g), list(g.edges()), b, gm))
    sample = [a for n in sorted(by_n) for a in by_n[n][:4]]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(procs) as pool:
        res = pool.map(_qaoa_cpu_value, sample, chunksize=1)
    wall = time.perf_counter() - t0
    mean_t = {}
    for n, dt in res:
        mean_t.setdefault(n, []).append(dt)
    mean_t = {n: float(np.mean(v)) for n, v in mean_t.items()}
    est = sum(mean_t[len(g)] for g, _, _ in inst) / procs
    out["config3_all_cores"] = {
        "processes": procs, "sample": "%d instances (4 per n=10..22), wall %.2f s"
        % (len(sample), wall),
        "sample_instances_per_s": round(len(sample) / wall, 2),
        "estimated_512_batch_instances_per_s": round(512 / est, 2)}
    inst = regular_instances(64, 16, 24, 8, seed=12)
    pick = [i for i, (g, _, _) in enumerate(inst) if len(g) == 16][:2]
    torch.set_num_threads(procs)
    t0 = time.perf_counter()
    for i in pick:
        g, b, gm = inst[i]
        bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
        gt = torch.tensor(gm, dtype=torch.float64, requires_grad=True)
        O.qaoa_value_torch(16, list(g.edges()), bt, gt).backward()
    t = time.perf_counter() - t0
    out["config4_torch_autograd_cpu"] = {
        "threads": procs, "sample": "%d instances at n=16, p=8" % len(pick),
        "value_grad_per_s": round(len(pick) / t, 3)}
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
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 12-qubit, GOPS and extra CPU sections")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain device copy of the state (4 GiB read + 4 GiB "
                         "written at 28 qubits) to scale HBM byte counters against")
    ap.add_argument("--only", choices=["qaoa", "subops"],
                    help="qaoa: run only the batched QAOA launches (for profilers) and print "
                         "their timings; subops: time one-pass plans per sub-op kind")
    ap.add_argument("--pass-times", action="store_true",
                    help="with --only qaoa: also report each global pass's span and busy "
                         "times from the kernel's wall-clock stamps")
    args = ap.parse_args()
    if args.only == "subops":
        print(json.dumps({"subop_costs": subop_costs()}), flush=True)
        return
    if args.only == "qaoa":
        res = {"qaoa_batches": qaoa_batches_only(max(1, args.steps), args.pass_times)}
        if args.copy_calibration:
            import torch
            a = torch.zeros(2 ** 26, dtype=torch.complex128, device="cuda")   # 1 GiB
            b = torch.empty_like(a)
            b.copy_(a)
            torch.cuda.synchronize()
            res["copy_calibration_bytes"] = 2 ** 30
        print(json.dumps(res), flush=True)
        return
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
    fp64 = None
    if plan is not None and kname == "tile_pass":
        # the same launches against the FP64 peak: the plan's algorithmic
        # floating-point work (sub-op counts, fusion.HostPlan.flops)
        fl = plan.flops()
        per_launch = sum(fl) / len(fl)
        fp64 = {"flops_per_launch": round(per_launch), "achieved_TFLOPs":
                round(per_launch / (avg_ms * 1e-3) / 1e12, 2), "peak_TFLOPs": FP64_PEAK_TFLOPS,
                "frac": round(per_launch / (avg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)}

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
        "tile_pass_fp64": fp64,
    }
    if plan is not None:
        out["config"].update(plan.stats())
    # correctness of the timed path at this size: the fused plan and the
    # unfused per-gate path, each once from |0...0>, compared amplitude by
    # amplitude (the unfused run is also the unfused throughput figure)
    note("timed %d steps: %.3f ms per step; checking fused vs unfused" % (args.steps, ms))
    if plan is not None and not args.no_unfused:
        work.zero_()
        work[0] = 1.0
        plan.run_inplace(work)
        ref = torch.zeros_like(work)
        ref[0] = 1.0
        unf = fusion.KernelTimer()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fusion.apply_gates_inplace(gates, qubits, ref, unf)
        torch.cuda.synchronize()
        dtu = time.perf_counter() - t0
        err = float((work - ref).abs().max().item())
        norm = float(qf.backend.norm(work))
        un_name, un_ms, _ = unf.dominant()
        out["unfused"] = {"gates_per_s": round(len(gates) / dtu, 2),
                          "dominant_kernel": un_name, "avg_ms": round(un_ms, 4),
                          "achieved_GBs": round(bytes_per_pass / (un_ms * 1e-3) / 1e9, 1)}
        out["check"] = {"max_abs_diff_fused_vs_unfused": err, "norm_minus_1": norm - 1.0,
                        "tolerance": 1e-10, "ok": bool(err <= 1e-10 and abs(norm - 1.0) <= 1e-10)}
        del ref
        torch.cuda.empty_cache()
    if args.copy_calibration:
        spare = torch.empty_like(work)
        spare.copy_(work)
        torch.cuda.synchronize()
        del spare
    if not args.no_mfma:
        note("matrix-core block plans")
        out["fused_mfma_blocks"] = fused_block_section(gates, qubits, work)
    if not args.no_extra:
        note("12-qubit circuit and GOPS")
        out["config1"] = config1_section()
        out["gops"] = gops_section()
    if not args.no_qaoa:
        del state
        note("QAOA batches and the plain path on the same instances")
        out["qaoa"] = qaoa_section(max(2, args.steps))
    if not args.no_density:
        note("density evolution")
        out["density"] = density_section()
    if not args.no_cpu and args.cpu_gates > 0:
        del work
        torch.cuda.empty_cache()
        note("CPU baselines")
        out["cpu_baseline"] = cpu_baseline(desc, n, args.cpu_gates)
        if not args.no_extra:
            out["cpu_baselines_extra"] = cpu_extra_baselines(
                random_mixed_circuit(20, DEPTH, seed=SEED), random_mixed_circuit(24, DEPTH, seed=SEED))
    print(json.dumps(out), flush=True)
    if "check" in out and not out["check"]["ok"]:
        print("bench.py: fused and unfused states disagree (%s)" % out["check"], file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()

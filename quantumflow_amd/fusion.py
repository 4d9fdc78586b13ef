"""Gate fusion and LDS-tiled execution plans, plus launch timing helpers."""

from collections import defaultdict

import torch

from . import backend as bk


class KernelTimer:
    """HIP-event timing of individual launches on torch's current stream
    (the stream every qfx call is enqueued on)."""

    def __init__(self):
        self.events = defaultdict(list)

    def start(self, name):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return name, e

    def stop(self, tok):
        name, e0 = tok
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.events[name].append((e0, e1))

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for name, pairs in self.events.items():
            ts = [a.elapsed_time(b) for a, b in pairs]
            out[name] = (sum(ts), len(ts))
        return out

    def dominant(self):
        """(kernel name, average ms per launch, launches) of the kernel class
        with the largest total time."""
        s = self.summary()
        if not s:
            return None, float('nan'), 0
        name = max(s, key=lambda k: s[k][0])
        tot, cnt = s[name]
        return name, tot / cnt, cnt


def kernel_class(kind):
    return '%s_k%d' % (kind.kind, kind.k)


def apply_gates_inplace(gates, qubits, work, prof=None):
    """Unfused path on a flat device state ``work``: one in-place kernel per
    gate (diagonal/monomial gates touch only the rows they change)."""
    qubits = tuple(qubits)
    n = len(qubits)
    for g in gates:
        bits = [n - 1 - qubits.index(q) for q in g.qubits]
        kind = g.kind
        if prof is not None:
            tok = prof.start(kernel_class(kind))
        bk.apply_gate(work, work, n, bits, kind, inplace_sparse=True)
        if prof is not None:
            prof.stop(tok)
    return work


def run_unfused_profiled(gates, qubits, work, prof):
    return apply_gates_inplace(gates, qubits, work, prof)


def plan_for(gates, qubits):
    """Fused execution plan for ``gates`` on a state with these qubits, or
    None when the unfused path should run."""
    return None

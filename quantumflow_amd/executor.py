"""Circuit executors: run a whole gate list on one working copy of the state.

``run_gates_inplace`` is the unfused path: one kernel per gate, in place on
a private copy of the input, with diagonal and monomial gates touching only
the amplitude rows they change.  ``run_gates`` picks the fused, LDS-tiled
executor when it is available for the gate list (see quantumflow_amd.fusion).
"""

from . import backend as bk
from .states import State


def gate_bits(gate, qubits):
    """Index bit positions (n-1-axis) of a gate's qubits in a state."""
    n = len(qubits)
    return [n - 1 - qubits.index(q) for q in gate.qubits]


def run_gates_inplace(gates, ket: State, work=None) -> State:
    """Apply ``gates`` in order, one in-place kernel each, to a copy of
    ``ket`` (or into ``work``, a flat complex128 device tensor the size of
    the state, overwritten)."""
    qubits = tuple(ket.qubits)
    n = len(qubits)
    src = bk.as_device(ket.tensor)
    if work is None:
        work = src.clone()
    elif work.data_ptr() != src.data_ptr():
        work.copy_(src)
    from .fusion import apply_gates_inplace
    apply_gates_inplace(gates, qubits, work)
    return State(work.view([2] * n), qubits, ket.memory)


def run_gates(gates, ket: State) -> State:
    from . import fusion
    plan = fusion.plan_for(gates, ket.qubits)
    if plan is None:
        return run_gates_inplace(gates, ket)
    return plan.run(ket)

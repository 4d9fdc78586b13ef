"""Quantum circuits (quantumflow/circuits.py:35-372).

``Circuit.run`` applies its elements in order to a state (circuits.py:87-99).
For plain Gates on a state of at least ``FUSED_MIN_QUBITS`` qubits it hands
the whole gate list to the fused executor (quantumflow_amd.executor), which
plans LDS-tiled passes and fused dense blocks and runs them on the GPU;
``run(ket, fused=False)`` forces the plain path, one kernel per gate.
"""

from collections import defaultdict
from itertools import chain
from math import pi
from typing import Dict, Iterable, Iterator, Sequence, Type

from . import config
from .gates import control_gate, identity_gate
from .ops import Gate, Operation
from .qubits import Qubit, Qubits
from .states import State, zero_state
from .stdgates import CCNOT, CNOT, CPHASE, H, SWAP, T, TY, TZ, X

__all__ = ['Circuit', 'count_operations', 'map_gate', 'qft_circuit',
           'reversal_circuit', 'control_circuit', 'ccnot_circuit', 'zyz_circuit',
           'phase_estimation_circuit', 'addition_circuit', 'ghz_circuit']

FUSED_MIN_QUBITS = 14


class Circuit(Operation):
    """A sequence of operations (circuits.py:35-136)."""

    def __init__(self, elements: Iterable[Operation] = None) -> None:
        self.elements = list(elements) if elements is not None else []

    def add(self, other: 'Circuit') -> 'Circuit':
        return Circuit(self.elements + other.elements)

    def extend(self, other: Operation) -> None:
        if isinstance(other, Circuit):
            self.elements.extend(other.elements)
        else:
            self.elements.extend([other])

    def __add__(self, other: 'Circuit') -> 'Circuit':
        return self.add(other)

    def __iadd__(self, other: Operation) -> 'Circuit':
        self.extend(other)
        return self

    def __iter__(self) -> Iterator[Operation]:
        return self.elements.__iter__()

    def size(self) -> int:
        return len(self.elements)

    @property
    def qubits(self) -> Qubits:
        """Sorted unique qubits; TypeError if unsortable (circuits.py:75-85)."""
        qbs = [q for elem in self.elements for q in elem.qubits]
        qbs = list(set(qbs))
        qbs = sorted(qbs)
        return tuple(qbs)

    def flat_elements(self) -> list:
        """All elements in order with nested circuits flattened."""
        out = []
        for elem in self.elements:
            if isinstance(elem, Circuit):
                out.extend(elem.flat_elements())
            else:
                out.append(elem)
        return out

    def flat_gates(self):
        """All elements as plain Gates in order, or None if any element is
        not a Gate (nested circuits are flattened)."""
        out = self.flat_elements()
        if any(not isinstance(elem, Gate) for elem in out):
            return None
        return out

    def evolve(self, rho=None):
        """Apply every element to a Density in order (circuits.py:102-109);
        default |0...0><0...0| on the circuit's qubits."""
        if rho is None:
            rho = zero_state(qubits=self.qubits).asdensity()
        for elem in self.elements:
            rho = elem.evolve(rho)
        return rho

    def run(self, ket: State = None, fused: bool = None) -> State:
        """Apply the circuit to ``ket`` (default |0...0> on the sorted
        qubits).  ``fused=None`` picks the fused executor for circuits on
        >= FUSED_MIN_QUBITS qubits when QUANTUMFLOW_FUSION is on: each
        maximal run of plain gates between other operations (measurements,
        resets, classical control) goes through it; the others run one by
        one in order."""
        if ket is None:
            ket = zero_state(qubits=self.qubits)
        if fused is None:
            fused = config.FUSION and ket.qubit_nb >= FUSED_MIN_QUBITS
        if not fused:
            for elem in self.elements:
                ket = elem.run(ket)
            return ket
        from .executor import run_gates
        segment = []
        for elem in self.flat_elements() + [None]:
            if isinstance(elem, Gate):
                segment.append(elem)
                continue
            if len(segment) > 1:
                ket = run_gates(segment, ket)
            elif segment:
                ket = segment[0].run(ket)
            segment = []
            if elem is not None:
                ket = elem.run(ket)
        return ket

    def asgate(self) -> Gate:
        """The circuit's unitary, composed on the GPU (circuits.py:111-118)."""
        gate = identity_gate(self.qubits)
        for elem in self.elements:
            gate = elem.asgate() @ gate
        return gate

    def aschannel(self):
        """The circuit's superoperator: the identity channel on its qubits
        composed with each element's channel in order (circuits.py:122-127);
        each composition is one gate-kernel pass over the superoperator."""
        chan = identity_gate(self.qubits).aschannel()
        for elem in self.elements:
            chan = elem.aschannel() @ chan
        return chan

    @property
    def H(self) -> 'Circuit':
        return Circuit([elem.H for elem in self.elements[::-1]])

    def __str__(self) -> str:
        return '\n'.join([str(elem) for elem in self.elements])


def count_operations(elements: Iterable[Operation]) -> Dict[Type[Operation], int]:
    """Count operations by type (circuits.py:139-149)."""
    op_count: Dict[Type[Operation], int] = defaultdict(int)
    for elem in elements:
        op_count[type(elem)] += 1
    return dict(op_count)


def map_gate(gate: Gate, args: Sequence[Qubits]) -> Circuit:
    """The same gate on each qubit tuple (circuits.py:152-167)."""
    circ = Circuit()
    for qubits in args:
        circ += gate.relabel(qubits)
    return circ


def qft_circuit(qubits: Qubits) -> Circuit:
    """Quantum Fourier transform (circuits.py:170-185)."""
    N = len(qubits)
    circ = Circuit()
    for n0 in range(N):
        q0 = qubits[n0]
        circ += H(q0)
        for n1 in range(n0 + 1, N):
            circ += CPHASE(pi / 2 ** (n1 - n0), qubits[n1], q0)
    circ.extend(reversal_circuit(qubits))
    return circ


def reversal_circuit(qubits: Qubits) -> Circuit:
    """Reverse the qubit order with SWAPs (circuits.py:188-194)."""
    N = len(qubits)
    circ = Circuit()
    for n in range(N // 2):
        circ += SWAP(qubits[n], qubits[N - 1 - n])
    return circ


def control_circuit(controls: Qubits, gate: Gate) -> Circuit:
    """Multi-controlled gate, Barenco et al. 1995 sec. 7.2 (circuits.py:197-225)."""
    circ = Circuit()
    if len(controls) == 1:
        q0 = controls[0]
        if isinstance(gate, X):
            circ += CNOT(q0, gate.qubits[0])
        else:
            circ += control_gate(q0, gate)
    else:
        circ += control_circuit(controls[-1:], gate ** 0.5)
        circ += control_circuit(controls[0:-1], X(controls[-1]))
        circ += control_circuit(controls[-1:], gate ** -0.5)
        circ += control_circuit(controls[0:-1], X(controls[-1]))
        circ += control_circuit(controls[0:-1], gate ** 0.5)
    return circ


def ccnot_circuit(qubits: Qubits) -> Circuit:
    """Toffoli from CNOT, H and T (Nielsen & Chuang; circuits.py:228-259)."""
    if len(qubits) != 3:
        raise ValueError('Expected 3 qubits')
    q0, q1, q2 = qubits
    circ = Circuit()
    for g in [H(q2), CNOT(q1, q2), T(q2).H, CNOT(q0, q2), T(q2), CNOT(q1, q2),
              T(q2).H, CNOT(q0, q2), T(q1), T(q2), H(q2), CNOT(q0, q1), T(q0),
              T(q1).H, CNOT(q0, q1)]:
        circ += g
    return circ


def zyz_circuit(t0: float, t1: float, t2: float, q0: Qubit) -> Circuit:
    """TZ(t0), TY(t1), TZ(t2) (circuits.py:262-268)."""
    circ = Circuit()
    circ += TZ(t0, q0)
    circ += TY(t1, q0)
    circ += TZ(t2, q0)
    return circ


def phase_estimation_circuit(gate: Gate, outputs: Qubits) -> Circuit:
    """Quantum phase estimation (circuits.py:271-303)."""
    circ = Circuit()
    circ += map_gate(H(), list(zip(outputs)))
    for cq in reversed(outputs):
        circ += control_gate(cq, gate)
        gate = gate @ gate
    circ += qft_circuit(outputs).H
    return circ


def addition_circuit(addend0: Qubits, addend1: Qubits, carry: Qubits) -> Circuit:
    """Cuccaro ripple-carry adder; result in addend1 (circuits.py:306-356)."""
    if len(addend0) != len(addend1):
        raise ValueError('Number of addend qubits must be equal')
    if len(carry) != 2:
        raise ValueError('Expected 2 carry qubits')

    def _maj(qubits: Qubits) -> Circuit:
        q0, q1, q2 = qubits
        return Circuit([CNOT(q2, q1), CNOT(q2, q0), CCNOT(q0, q1, q2)])

    def _uma(qubits: Qubits) -> Circuit:
        q0, q1, q2 = qubits
        return Circuit([CCNOT(q0, q1, q2), CNOT(q2, q0), CNOT(q0, q1)])

    qubits = [carry[0]] + list(chain.from_iterable(
        zip(reversed(addend1), reversed(addend0)))) + [carry[1]]
    circ = Circuit()
    for n in range(0, len(qubits) - 3, 2):
        circ += _maj(qubits[n:n + 3])
    circ += CNOT(qubits[-2], qubits[-1])
    for n in reversed(range(0, len(qubits) - 3, 2)):
        circ += _uma(qubits[n:n + 3])
    return circ


def ghz_circuit(qubits: Qubits) -> Circuit:
    """H then a CNOT ladder (circuits.py:359-371)."""
    circ = Circuit()
    circ += H(qubits[0])
    for q0 in range(0, len(qubits) - 1):
        circ += CNOT(qubits[q0], qubits[q0 + 1])
    return circ

"""Measurement and reset operations on the device state
(quantumflow/stdops.py:38-126).

``Measure`` reads P(|0>) of one qubit with a single reduction kernel
(``qfx_qubit_probs``), rolls numpy's global generator on the host exactly as
the reference does (stdops.py:56), and collapses and renormalises the state in
one diagonal-gate pass: diag(1/sqrt(p0), 0) or diag(0, 1/sqrt(p1)) on the
measured qubit, which is P0/P1 followed by ``normalize`` (stdops.py:57,61)
fused.  ``Reset`` applies [[1, 1], [0, 0]] to each qubit with the gate kernel
and normalises (stdops.py:100-111).  ``Barrier``, ``If`` and ``Move`` are the
control and classical-memory operations the reference's own measurement tests
use (stdops.py:129-170, :279-282).
"""

import operator
from math import sqrt
from typing import Any, Callable, Hashable

import numpy as np

from . import backend as bk
from .ops import Operation
from .qubits import Qubit, Qubits
from .states import State

__all__ = ['Measure', 'Reset', 'Barrier', 'If', 'Move', 'Neg', 'Not', 'And', 'Ior', 'Or',
           'Xor', 'Add', 'Sub', 'Mul', 'Div', 'Exchange', 'EQ', 'GT', 'GE', 'LT', 'LE', 'NE']

_RESET = np.array([[1, 1], [0, 0]], dtype=np.complex128)


class Measure(Operation):
    """Measure a qubit in the computational basis and, if ``cbit`` is given,
    store the outcome in the state's classical memory (stdops.py:38-65)."""

    def __init__(self, qubit: Qubit, cbit: Hashable = None) -> None:
        self.qubit = qubit
        self.cbit = cbit

    def quil(self) -> str:
        if self.cbit is not None:
            return '{} {} {}'.format(self.name.upper(), self.qubit, self.cbit)
        return '{} {}'.format(self.name.upper(), self.qubit)

    @property
    def qubits(self) -> Qubits:
        return [self.qubit]

    def run(self, ket: State) -> State:
        axis = ket.qubits.index(self.qubit)
        p = bk.evaluate(bk.qubit_probs(ket.tensor, axis))
        prob_zero, prob_one = float(p[0]), float(p[1])
        if np.random.random() < prob_zero:
            bit, d = 0, [1.0 / sqrt(prob_zero), 0.0]
        else:
            # the reference divides by sqrt(norm) here too; a zero branch
            # probability gives inf/nan amplitudes there as well
            bit, d = 1, [0.0, 1.0 / sqrt(prob_one) if prob_one > 0 else float('inf')]
        tensor = bk.tensormul(bk.classify(np.diag(np.asarray(d, dtype=np.complex128))),
                              ket.tensor, [axis])
        ket = State(tensor, ket.qubits, ket.memory)
        if self.cbit is not None:
            ket = ket.update({self.cbit: bit})
        return ket

    def evolve(self, rho):
        """Measure a Density (stdops.py:67-82): P0 rho P0 and P1 rho P1 are
        one gate-kernel pass each.  As in the reference, the outcome 0 is
        drawn with the value of the projected density's norm() (its squared
        Frobenius norm), and the kept projection is divided by its trace."""
        from .gates import P0, P1
        r0 = P0(self.qubit).evolve(rho)
        prob_zero = float(np.abs(bk.evaluate(r0.norm())))
        if np.random.random() < prob_zero:
            bit, rho = 0, r0.normalize()
        else:
            bit, rho = 1, P1(self.qubit).evolve(rho).normalize()
        if self.cbit is not None:
            rho = rho.update({self.cbit: bit})
        return rho


class Reset(Operation):
    """Reset qubits (all of the state's when none are given) to |0>
    irrespective of the initial state (stdops.py:85-111)."""

    def __init__(self, *qubits: Qubit) -> None:
        self._qubits = tuple(qubits)

    @property
    def H(self) -> 'Reset':
        return self  # Hermitian

    def quil(self) -> str:
        if self.qubits:
            return 'RESET ' + ' '.join(str(q) for q in self.qubits)
        return 'RESET'

    def run(self, ket: State) -> State:
        qubits = self.qubits if self.qubits else ket.qubits
        kind = bk.classify(_RESET)
        tensor = ket.tensor
        for q in qubits:
            tensor = bk.tensormul(kind, tensor, [ket.qubits.index(q)])
        return State(tensor, ket.qubits, ket.memory).normalize()

    def evolve(self, rho):
        """Not defined for densities, as in the reference (stdops.py:113-115)."""
        raise TypeError('Not yet implemented')


class Barrier(Operation):
    """No-op that marks a scheduling barrier (stdops.py:129-148)."""

    def __init__(self, *qubits: Qubit) -> None:
        self._qubits = tuple(qubits)

    @property
    def H(self) -> 'Barrier':
        return self

    def quil(self) -> str:
        return self.name.upper() + ' ' + ' '.join(str(q) for q in self.qubits)

    def run(self, ket: State) -> State:
        return ket

    def evolve(self, rho):
        return rho


class If(Operation):
    """Apply ``elem`` when classical memory at ``condition`` equals
    ``value`` (stdops.py:151-170)."""

    def __init__(self, elem: Operation, condition: Hashable, value: bool = True) -> None:
        self.element = elem
        self.value = value
        self.condition = condition

    @property
    def qubits(self) -> Qubits:
        return self.element.qubits

    def run(self, ket: State) -> State:
        if ket.memory[self.condition] == self.value:
            ket = self.element.run(ket)
        return ket

    def evolve(self, rho):
        """stdops.py:165-168."""
        if rho.memory[self.condition] == self.value:
            rho = self.element.evolve(rho)
        return rho


class Move(Operation):
    """Copy a value (or another memory cell's value) into classical memory
    (stdops.py:279-282; source lookup as BinaryOP._source, :201-204)."""

    def __init__(self, target: Hashable, source: Any) -> None:
        self.target = target
        self.source = source

    def _value(self, ket: State):
        # a number is a constant; anything else names a memory cell
        if isinstance(self.source, (bool, int, float, complex, np.number)):
            return self.source
        return ket.memory[self.source]

    def quil(self) -> str:
        return 'MOVE {} {}'.format(self.target, self.source)

    def run(self, ket: State) -> State:
        return ket.update({self.target: self._value(ket)})

    def evolve(self, rho):
        return self.run(rho)



def _is_number(x) -> bool:
    return isinstance(x, (bool, int, float, complex, np.number))


class Neg(Operation):
    """Negate a value in classical memory (stdops.py:173-182)."""

    def __init__(self, target: Hashable) -> None:
        self.target = target

    def quil(self) -> str:
        return '{} {}'.format(self.name.upper(), self.target)

    def run(self, ket: State) -> State:
        return ket.update({self.target: -ket.memory[self.target]})

    def evolve(self, rho):
        return self.run(rho)


class Not(Operation):
    """Logical NOT of a classical bit (stdops.py:185-194)."""

    def __init__(self, target: Hashable) -> None:
        self.target = target

    def quil(self) -> str:
        return '{} {}'.format(self.name.upper(), self.target)

    def run(self, ket: State) -> State:
        return ket.update({self.target: int(not ket.memory[self.target])})

    def evolve(self, rho):
        return self.run(rho)


class _BinaryOp(Operation):
    """target <- op(target, source), source an address or a constant
    (stdops.py:199-229)."""
    _op: Callable = None

    def __init__(self, target: Hashable, source: Any) -> None:
        self.target = target
        self.source = source

    def _source(self, ket):
        return self.source if _is_number(self.source) else ket.memory[self.source]

    def quil(self) -> str:
        return '{} {} {}'.format(self.name.upper(), self.target, self.source)

    def run(self, ket: State) -> State:
        return ket.update({self.target: type(self)._op(ket.memory[self.target],
                                                       self._source(ket))})

    def evolve(self, rho):
        return self.run(rho)


class And(_BinaryOp):
    """Logical AND into the target (stdops.py:232-234)."""
    _op = operator.and_


class Ior(_BinaryOp):
    """Inclusive OR into the target (stdops.py:237-240)."""
    _op = operator.or_


class Or(_BinaryOp):
    """Inclusive OR (Quil's deprecated name for IOR, stdops.py:243-246)."""
    _op = operator.or_


class Xor(_BinaryOp):
    """Exclusive OR into the target (stdops.py:249-253)."""
    _op = operator.xor


class Add(_BinaryOp):
    """target + source (stdops.py:256-258)."""
    _op = operator.add


class Sub(_BinaryOp):
    """target - source (stdops.py:261-263)."""
    _op = operator.sub


class Mul(_BinaryOp):
    """target * source (stdops.py:266-268)."""
    _op = operator.mul


class Div(_BinaryOp):
    """target / source, true division (stdops.py:271-273)."""
    _op = operator.truediv


class Exchange(_BinaryOp):
    """Swap two classical cells (stdops.py:285-290)."""

    def run(self, ket: State) -> State:
        mem = ket.memory
        return ket.update({self.target: mem[self.source], self.source: mem[self.target]})


class _Comparison(Operation):
    """target <- op(left, right) on two addresses (stdops.py:295-311)."""
    _op: Callable = None

    def __init__(self, target: Hashable, left: Hashable, right: Hashable) -> None:
        self.target = target
        self.left = left
        self.right = right

    def quil(self) -> str:
        return '{} {} {} {}'.format(self.name, self.target, self.left, self.right)

    def run(self, ket: State) -> State:
        mem = ket.memory
        return ket.update({self.target: type(self)._op(mem[self.left], mem[self.right])})

    def evolve(self, rho):
        return self.run(rho)


class EQ(_Comparison):
    """left == right (stdops.py:314-316)."""
    _op = operator.eq


class GT(_Comparison):
    """left > right (stdops.py:319-321)."""
    _op = operator.gt


class GE(_Comparison):
    """left >= right (stdops.py:324-326)."""
    _op = operator.ge


class LT(_Comparison):
    """left < right (stdops.py:329-331)."""
    _op = operator.lt


class LE(_Comparison):
    """left <= right (stdops.py:334-336)."""
    _op = operator.le


class NE(_Comparison):
    """left != right (stdops.py:339-341)."""
    _op = operator.ne

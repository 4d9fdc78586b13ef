"""Gate utilities (quantumflow/gates.py:1-168).

Small-matrix gate algebra (Kronecker joins, controls, projectors) is done on
the host, where the gate matrices are built; applying gates to states and
composing gates with ``@`` run on the GPU.
"""

from functools import reduce
from typing import TextIO, Union

import numpy as np
import scipy.stats

from .config import TOLERANCE
from .ops import Gate
from .qubits import Qubit, Qubits, qubits_count_tuple, asarray

__all__ = ['I', 'identity_gate', 'random_gate', 'join_gates', 'control_gate',
           'conditional_gate', 'P0', 'P1', 'almost_unitary', 'almost_identity',
           'almost_hermitian', 'print_gate']


class I(Gate):  # noqa: E742
    """The identity gate on any number of qubits (gates.py:33-54)."""

    def __init__(self, *qubits: Qubit) -> None:
        if not qubits:
            qubits = (0,)
        N = len(qubits)
        super().__init__(np.eye(2 ** N), qubits=qubits)

    @property
    def H(self) -> Gate:
        return self

    def __pow__(self, t: float) -> Gate:
        return self


def identity_gate(qubits: Union[int, Qubits]) -> Gate:
    """The K-qubit identity gate (gates.py:57-60)."""
    _, qubits = qubits_count_tuple(qubits)
    return I(*qubits)


def join_gates(*gates: Gate) -> Gate:
    """Direct product of gates on disjoint qubits (gates.py:63-68)."""
    qubits = tuple(q for g in gates for q in g.qubits)
    if len(set(qubits)) != len(qubits):
        raise ValueError('Overlapping qubits')
    m = reduce(np.kron, [g.matrix for g in gates])
    return Gate(m, qubits)


def control_gate(control: Qubit, gate: Gate) -> Gate:
    """Controlled gate with the control prepended (gates.py:71-83)."""
    if control in gate.qubits:
        raise ValueError('Gate and control qubits overlap')
    qubits = [control, *gate.qubits]
    d = 2 ** gate.qubit_nb
    m = np.zeros((2 * d, 2 * d), dtype=np.complex128)
    m[:d, :d] = np.eye(d)
    m[d:, d:] = gate.matrix
    return Gate(m, qubits)


def conditional_gate(control: Qubit, gate0: Gate, gate1: Gate) -> Gate:
    """gate0 if the control is 0 else gate1 (gates.py:86-94)."""
    assert gate0.qubits == gate1.qubits
    d = 2 ** gate0.qubit_nb
    m = np.zeros((2 * d, 2 * d), dtype=np.complex128)
    m[:d, :d] = gate0.matrix
    m[d:, d:] = gate1.matrix
    return Gate(m, [control, *gate0.qubits])


def almost_unitary(gate: Gate) -> bool:
    m = gate.matrix
    return np.allclose(m @ m.conj().T, np.eye(m.shape[0]), atol=TOLERANCE)


def almost_identity(gate: Gate) -> bool:
    return np.allclose(gate.matrix, np.eye(gate.matrix.shape[0]))


def almost_hermitian(gate: Gate) -> bool:
    return np.allclose(gate.matrix, gate.matrix.conj().T)


def print_gate(gate: Gate, ndigits: int = 2, file: TextIO = None) -> None:
    N = gate.qubit_nb
    t = gate.matrix.reshape([2] * (2 * N))
    lines = []
    for index, amplitude in np.ndenumerate(t):
        ket = "".join([str(n) for n in index[0:N]])
        bra = "".join([str(index[n]) for n in range(N, 2 * N)])
        if round(abs(amplitude) ** 2, ndigits) > 0.0:
            lines.append('{} -> {} : {}'.format(bra, ket, amplitude))
    lines.sort(key=lambda x: int(x[0:N]))
    print('\n'.join(lines), file=file)


class P0(Gate):
    """Projector onto |0> (gates.py:137-145)."""

    def __init__(self, q0: Qubit = 0) -> None:
        super().__init__([[1, 0], [0, 0]], qubits=[q0])


class P1(Gate):
    """Projector onto |1> (gates.py:148-156)."""

    def __init__(self, q0: Qubit = 0) -> None:
        super().__init__([[0, 0], [0, 1]], qubits=[q0])


def random_gate(qubits: Union[int, Qubits]) -> Gate:
    """Haar-random unitary on K qubits (gates.py:159-168)."""
    N, qubits = qubits_count_tuple(qubits)
    unitary = scipy.stats.unitary_group.rvs(2 ** N)
    return Gate(unitary, qubits=qubits, name='RAND{}'.format(N))

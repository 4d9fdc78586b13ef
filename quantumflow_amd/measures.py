"""Distances between states and gates (quantumflow/measures.py:32-53,
183-194): fidelity and Fubini-Study angles, computed with the GPU inner
product."""

from . import backend as bk
from .config import TOLERANCE
from .ops import Gate
from .qubits import fubini_study_angle, vectors_close
from .states import State

__all__ = ['state_fidelity', 'state_angle', 'states_close', 'gate_angle',
           'gates_close']


def state_fidelity(state0: State, state1: State) -> bk.BKTensor:
    """|<psi0|psi1>|^2 (measures.py:32-36)."""
    assert state0.qubits == state1.qubits
    t = bk.inner(state0.tensor, state1.tensor)
    return bk.absolute(t) ** 2


def state_angle(ket0: State, ket1: State) -> bk.BKTensor:
    """Fubini-Study angle between states (measures.py:39-44)."""
    return fubini_study_angle(ket0.vec, ket1.vec)


def states_close(state0: State, state1: State, tolerance: float = TOLERANCE) -> bool:
    """True if the states are close in projective Hilbert space (measures.py:47-53)."""
    return vectors_close(state0.vec, state1.vec, tolerance)


def gate_angle(gate0: Gate, gate1: Gate) -> bk.BKTensor:
    """Fubini-Study angle between gates (measures.py:183-185)."""
    return fubini_study_angle(gate0.vec, gate1.vec)


def gates_close(gate0: Gate, gate1: Gate, tolerance: float = TOLERANCE) -> bool:
    """True if the gates are close up to phase (measures.py:188-194)."""
    return vectors_close(gate0.vec, gate1.vec, tolerance)

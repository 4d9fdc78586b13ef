"""Pure quantum states (quantumflow/states.py:29-259).

A State holds a complex128 device tensor of shape [2]*N (16*2^N bytes of
HBM) and qubit labels.  Readouts run as single HIP kernels:
``probabilities`` (|psi|^2), ``expectation`` (sum diag |psi|^2),
``expectation_z`` (all <Z_i> in one pass), ``norm`` (|<psi|psi>|).
"""

from collections import defaultdict
from functools import reduce
from math import sqrt
from typing import Any, Dict, Sequence, TextIO, Union

import numpy as np

from . import backend as bk
from .qubits import Qubits, QubitVector, outer_product, qubits_count_tuple

__all__ = ['State', 'ghz_state', 'join_states', 'print_probabilities',
           'print_state', 'random_state', 'w_state', 'zero_state']


class State:
    """The quantum state of a collection of qubits (states.py:29-55)."""

    def __init__(self, tensor: bk.TensorLike, qubits: Qubits = None,
                 memory: Dict[Any, Any] = None) -> None:
        if qubits is None:
            tensor = bk.astensorproduct(tensor)
            qubits = range(bk.rank(tensor))
        self.vec = QubitVector(tensor, qubits)
        self._memory = memory if memory is not None else {}

    @property
    def tensor(self) -> bk.BKTensor:
        return self.vec.tensor

    @property
    def qubits(self) -> Qubits:
        return self.vec.qubits

    @property
    def qubit_nb(self) -> int:
        return self.vec.qubit_nb

    def norm(self) -> bk.BKTensor:
        """|<psi|psi>| (states.py:68-70, qubits.py:180-182)."""
        return self.vec.norm()

    @property
    def memory(self) -> dict:
        return defaultdict(int, self._memory)

    def asdensity(self):
        """|psi><psi| as a Density, formed on the device (states.py:162-165)."""
        from .channels import state_asdensity
        return state_asdensity(self)

    def update(self, memory: Dict[Any, Any]) -> 'State':
        mem = self.memory
        mem.update(memory)
        return State(self.tensor, self.qubits, mem)

    def relabel(self, qubits: Qubits) -> 'State':
        return State(self.vec.tensor, qubits, self._memory)

    def permute(self, qubits: Qubits) -> 'State':
        vec = self.vec.permute(qubits)
        return State(vec.tensor, vec.qubits, self._memory)

    def normalize(self) -> 'State':
        """Divide by sqrt(norm()) (states.py:108-111)."""
        tensor = self.tensor / bk.sqrt(self.norm())
        return State(tensor, self.qubits, self._memory)

    def probabilities(self) -> bk.BKTensor:
        """|psi|^2, float64 [2]*N on the device (states.py:113-119)."""
        return bk.probabilities(self.tensor)

    def sample(self, trials: int) -> np.ndarray:
        """Counts of each basis configuration over ``trials`` measurements,
        as a host int64 array [2]*N (states.py:121-129).  The draws run on
        the device (prefix sums of |psi|^2, inverse CDF per draw); the seed
        comes from numpy's global generator, so np.random.seed makes them
        reproducible."""
        counts = bk.sample_counts(self.tensor, trials, bk.draw_seed())
        return bk.evaluate(counts)

    def sample_indices(self, trials: int, seed: int = None) -> bk.BKTensor:
        """``trials`` flat basis indices drawn on the device, int64 device
        tensor (index bit N-1-a is axis a); ``seed`` defaults to a draw from
        numpy's global generator."""
        if seed is None:
            seed = bk.draw_seed()
        return bk.sample_indices(self.tensor, trials, seed)

    def expectation(self, diag_hermitian: bk.TensorLike, trials: int = None) -> bk.BKTensor:
        """Expectation of a diagonal observable (states.py:131-147): exact
        (one fused kernel) when trials is None, else the mean of the diagonal
        over ``trials`` device draws."""
        if trials is None:
            return bk.expectation(diag_hermitian, self.tensor)
        return bk.sample_expectation(diag_hermitian, self.tensor, trials, bk.draw_seed())

    def expectation_z(self) -> bk.BKTensor:
        """<Z_i> for every qubit, in qubit order, one pass over the state."""
        return bk.expectation_z(self.tensor)

    def measure(self) -> np.ndarray:
        """One computational-basis measurement as N bits in axis order
        (states.py:149-160), drawn on the device."""
        idx = int(bk.evaluate(self.sample_indices(1))[0])
        return np.array([(idx >> (self.qubit_nb - 1 - a)) & 1 for a in range(self.qubit_nb)])

    def __str__(self) -> str:
        state = self.vec.asarray()
        s = []
        count = 0
        for index, amplitude in np.ndenumerate(state):
            if not np.isclose(amplitude, 0.0):
                ket = '|' + ''.join([str(n) for n in index]) + '>'
                s.append('({c.real:0.04g}{c.imag:+0.04g}i) {k}'.format(c=amplitude, k=ket))
                count += 1
                if count > 64:
                    s.append('...')
                    break
        return ' + '.join(s)


def zero_state(qubits: Union[int, Qubits]) -> State:
    """|0...0> on N qubits, filled on the device (states.py:186-191)."""
    N, qubits = qubits_count_tuple(qubits)
    return State(bk.zero_state_tensor(N), qubits)


def w_state(qubits: Union[int, Qubits]) -> State:
    """W state (states.py:194-202)."""
    N, qubits = qubits_count_tuple(qubits)
    ket = np.zeros(shape=[2] * N)
    for n in range(N):
        idx = np.zeros(shape=N, dtype=int)
        idx[n] += 1
        ket[tuple(idx)] = 1 / sqrt(N)
    return State(ket, qubits)


def ghz_state(qubits: Union[int, Qubits]) -> State:
    """GHZ state (states.py:205-211)."""
    N, qubits = qubits_count_tuple(qubits)
    ket = np.zeros(shape=[2] * N)
    ket[(0,) * N] = 1 / sqrt(2)
    ket[(1,) * N] = 1 / sqrt(2)
    return State(ket, qubits)


def random_state(qubits: Union[int, Qubits]) -> State:
    """Haar-random state (states.py:214-219)."""
    N, qubits = qubits_count_tuple(qubits)
    ket = np.random.normal(size=([2] * N)) + 1j * np.random.normal(size=([2] * N))
    return State(ket, qubits).normalize()


def join_states(*states: State) -> State:
    """Tensor product of states on disjoint qubits (states.py:224-230)."""
    vectors = [ket.vec for ket in states]
    vec = reduce(outer_product, vectors)
    return State(vec.tensor, vec.qubits)


def print_state(state: State, file: TextIO = None) -> None:
    state_arr = state.vec.asarray()
    for index, amplitude in np.ndenumerate(state_arr):
        ket = "".join([str(n) for n in index])
        print(ket, ":", amplitude, file=file)


def print_probabilities(state: State, ndigits: int = 4, file: TextIO = None) -> None:
    prob = bk.evaluate(state.probabilities())
    for index, p in np.ndenumerate(prob):
        p = round(float(p), ndigits)
        if p == 0.0:
            continue
        ket = "".join([str(n) for n in index])
        print(ket, ":", p, file=file)

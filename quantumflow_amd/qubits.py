"""QubitVector: a tensor on labelled qubits (quantumflow/qubits.py:39-227).

A tensor of rank R on N qubits has shape [2]*(N*R); axis i <-> qubits[i]; the
flat index is big-endian in axis order (qubits.py:63-68).  Gates are rank 2
with axes [ket_out..., bra_in...] (qubits.py:70-80).
"""

from copy import copy
from typing import Any, Hashable, List, Sequence, Tuple, Union

import numpy as np

from . import backend as bk
from . import _native
from .config import TOLERANCE

__all__ = ['Qubit', 'Qubits', 'asarray', 'QubitVector', 'inner_product',
           'outer_product', 'fubini_study_angle', 'vectors_close']

Qubit = Hashable
Qubits = Sequence[Qubit]


def asarray(tensor: bk.BKTensor) -> bk.TensorLike:
    """Convert a backend tensor to a numpy array (device -> host)."""
    return bk.evaluate(tensor)


def permute_axes(tensor, perm):
    """New tensor whose axis a is old axis perm[a], done by the qubit
    relabelling kernel (a gather over index bits) for complex device tensors."""
    n = bk.rank(tensor)
    if list(perm) == list(range(n)):
        return tensor
    src = bk.as_device(tensor)
    out = src.new_empty(src.shape)
    src_bit = [0] * n
    for a in range(n):
        src_bit[n - 1 - a] = n - 1 - perm[a]
    _native.call('qfx_permute_qubits', bk.stream_handle(),
                 _native.ctypes.c_void_p(out.data_ptr()),
                 _native.ctypes.c_void_p(src.data_ptr()), n, _native.int_array(src_bit))
    return out.view([2] * n)


class QubitVector:
    """Container for a backend tensor in product form plus qubit labels
    (qubits.py:39-121).  ``rank`` is 1 for states, 2 for gates."""

    def __init__(self, tensor: bk.TensorLike, qubits: Qubits, rank: int = None) -> None:
        tensor = bk.astensorproduct(tensor)
        self.tensor = tensor
        N = len(qubits)
        if rank is None:
            rank = 1 if N == 0 else bk.rank(tensor) // N
        if rank not in [1, 2, 4, 8] or rank * N != bk.rank(tensor):
            raise ValueError('Incompatibility between tensor and qubits')
        self.qubits = tuple(qubits)
        self.qubit_nb = N
        self.rank = rank

    def __getitem__(self, key: Any) -> bk.BKTensor:
        return bk.getitem(self.tensor, key)

    def asarray(self) -> np.ndarray:
        return bk.evaluate(self.tensor)

    def flatten(self) -> bk.BKTensor:
        return bk.reshape(self.tensor, [2 ** self.qubit_nb] * self.rank)

    def relabel(self, qubits: Qubits) -> 'QubitVector':
        qubits = tuple(qubits)
        assert len(qubits) == self.qubit_nb
        vec = copy(self)
        vec.qubits = qubits
        return vec

    def permute(self, qubits: Qubits) -> 'QubitVector':
        """Reorder the axes to the label order ``qubits`` (qubits.py:145-161);
        ValueError if the labels differ."""
        qubits = tuple(qubits)
        if qubits == self.qubits:
            return self
        N = self.qubit_nb
        assert len(qubits) == N
        indices = [self.qubits.index(q) for q in qubits]
        perm: List[int] = []
        for rr in range(self.rank):
            perm += [rr * N + idx for idx in indices]
        return QubitVector(permute_axes(self.tensor, perm), qubits)

    @property
    def H(self) -> 'QubitVector':
        """Conjugate transpose of the (super)operator (qubits.py:163-178)."""
        N, R = self.qubit_nb, self.rank
        t = bk.reshape(self.tensor, [2 ** (N * R // 2)] * 2)
        t = bk.transpose(t)
        t = bk.reshape(t, [2] * R * N)
        return QubitVector(bk.conj(t), self.qubits)

    def norm(self) -> bk.BKTensor:
        """|<v|v>| -- the squared 2-norm (qubits.py:180-182)."""
        return bk.norm(self.tensor)

    def trace(self) -> bk.BKTensor:
        N, R = self.qubit_nb, self.rank
        if R == 1:
            raise ValueError('Cannot take trace of vector')
        return bk.trace(bk.reshape(self.tensor, [2 ** (N * R // 2)] * 2))


def inner_product(vec0: QubitVector, vec1: QubitVector) -> bk.BKTensor:
    """Hilbert-Schmidt inner product (qubits.py:233-245)."""
    if vec0.rank != vec1.rank or vec0.qubit_nb != vec1.qubit_nb:
        raise ValueError('Incompatibly vectors. Qubits and rank must match')
    vec1 = vec1.permute(vec0.qubits)
    return bk.inner(vec0.tensor, vec1.tensor)


def outer_product(vec0: QubitVector, vec1: QubitVector) -> QubitVector:
    """Direct product of qubit vectors on disjoint qubits (qubits.py:248-279)."""
    R, R1 = vec0.rank, vec1.rank
    N0, N1 = vec0.qubit_nb, vec1.qubit_nb
    if R != R1:
        raise ValueError('Incompatibly vectors. Rank must match')
    if not set(vec0.qubits).isdisjoint(vec1.qubits):
        raise ValueError('Overlapping qubits')
    qubits: Qubits = tuple(vec0.qubits) + tuple(vec1.qubits)
    tensor = bk.outer(vec0.tensor, vec1.tensor)
    tensor = bk.reshape(tensor, ([2 ** N0] * R) + ([2 ** N1] * R))
    perm = [idx for ij in zip(range(0, R), range(R, 2 * R)) for idx in ij]
    tensor = bk.transpose(tensor, perm)
    return QubitVector(tensor, qubits)


def fubini_study_angle(vec0: QubitVector, vec1: QubitVector) -> bk.BKTensor:
    """Fubini-Study angle between vectors (qubits.py:282-307)."""
    if vec0.rank != vec1.rank or vec0.qubit_nb != vec1.qubit_nb:
        raise ValueError('Incompatibly vectors. Qubits and rank must match')
    vec1 = vec1.permute(vec0.qubits)
    t0, t1 = vec0.tensor, vec1.tensor
    hs01 = bk.inner(t0, t1)
    hs00 = bk.inner(t0, t0)
    hs11 = bk.inner(t1, t1)
    ratio = bk.absolute(hs01) / bk.sqrt(bk.absolute(hs00 * hs11))
    ratio = bk.minimum(ratio, bk.fcast(1.))
    return bk.arccos(ratio)


def vectors_close(vec0: QubitVector, vec1: QubitVector,
                  tolerance: float = TOLERANCE) -> bool:
    """True if the vectors are close in projective Hilbert space
    (qubits.py:310-325)."""
    if vec0.rank != vec1.rank:
        return False
    if vec0.qubit_nb != vec1.qubit_nb:
        return False
    if set(vec0.qubits) ^ set(vec1.qubits):
        return False
    return bk.evaluate(fubini_study_angle(vec0, vec1)) <= tolerance


def qubits_count_tuple(qubits: Union[int, Qubits]) -> Tuple[int, Qubits]:
    """Unravel 'qubits: Union[int, Qubits]' arguments (qubits.py:329-333)."""
    if isinstance(qubits, int):
        return qubits, tuple(range(qubits))
    return len(qubits), qubits

"""Operations on states: the Gate (quantumflow/ops.py:35-259).

A Gate keeps its 2^K x 2^K operator either as a host numpy array (the
standard gates build theirs on the host with the reference's numpy
expressions, so entries agree bit for bit) or as a device tensor (gates
composed on the GPU by ``@``), converting lazily.  ``Gate.run`` passes a
host matrix to the kernel by value -- no host-to-device copy per gate.
"""

from abc import ABC
from copy import copy
from typing import Any, Dict

import numpy as np
import torch
from scipy.linalg import fractional_matrix_power as matpow

from . import backend as bk
from .qubits import Qubits, QubitVector
from .states import State

__all__ = ['Operation', 'Gate']


class Operation(ABC):
    """An operation on a qubit state (ops.py:35-88)."""

    _qubits: Qubits = ()

    @property
    def qubits(self) -> Qubits:
        return self._qubits

    @property
    def qubit_nb(self) -> int:
        return len(self.qubits)

    @property
    def name(self) -> str:
        return self.__class__.__name__.upper()

    def run(self, ket: State) -> State:
        raise NotImplementedError()

    def quil(self) -> str:
        raise NotImplementedError()

    def __str__(self) -> str:
        return self.quil()

    def asgate(self) -> 'Gate':
        raise NotImplementedError()

    @property
    def H(self) -> 'Operation':
        raise NotImplementedError()


def _to_host_matrix(tensor):
    if isinstance(tensor, torch.Tensor):
        arr = tensor.detach().cpu().numpy()
    else:
        arr = np.asarray(tensor)
    arr = np.asarray(arr, dtype=np.complex128)
    n = int(round(np.log2(arr.size)))
    if 2 ** n != arr.size or n % 2:
        raise ValueError('Incompatibility between tensor and qubits')
    k = n // 2
    return arr.reshape(2 ** k, 2 ** k)


class Gate(Operation):
    """A quantum logic gate on K qubits (ops.py:91-259)."""

    def __init__(self, tensor: bk.TensorLike, qubits: Qubits = None,
                 params: Dict[str, float] = None, name: str = None) -> None:
        self._host = None
        self._dev = None
        if isinstance(tensor, torch.Tensor) and tensor.device.type == 'cuda':
            dev = tensor.to(bk.CTYPE)
            n = int(round(np.log2(dev.numel())))
            if 2 ** n != dev.numel() or n % 2:
                raise ValueError('Incompatibility between tensor and qubits')
            k = n // 2
            self._dev = dev.reshape([2] * n) if n else dev
        else:
            self._host = _to_host_matrix(tensor)
            k = int(round(np.log2(self._host.shape[0])))
        if qubits is None:
            qubits = range(k)
        qubits = tuple(qubits)
        if len(qubits) != k:
            raise ValueError('Incompatibility between tensor and qubits')
        self._qubits = qubits
        self._kind = None
        self.params = params if params is not None else {}
        self._name = name if name is not None else self.__class__.__name__

    # ---- representations ----
    @property
    def name(self) -> str:
        return self._name

    @property
    def matrix(self) -> np.ndarray:
        """The operator as a host 2^K x 2^K complex128 array."""
        if self._host is None:
            self._host = _to_host_matrix(self._dev)
        return self._host

    @property
    def tensor(self) -> bk.BKTensor:
        """The gate tensor [2]*2K on the device (ops.py:128-131)."""
        if self._dev is None:
            self._dev = bk.astensorproduct(self._host)
        return self._dev

    @property
    def vec(self) -> QubitVector:
        return QubitVector(self.tensor, self.qubits)

    @property
    def kind(self):
        """Kernel class of this gate (dense / diagonal / monomial)."""
        if self._kind is None:
            self._kind = bk.classify(self.matrix)
        return self._kind

    @property
    def qubits(self) -> Qubits:
        return self._qubits

    @property
    def qubit_nb(self) -> int:
        return len(self._qubits)

    def relabel(self, qubits: Qubits) -> 'Gate':
        gate = copy(self)
        qubits = tuple(qubits)
        assert len(qubits) == self.qubit_nb
        gate._qubits = qubits
        return gate

    def permute(self, qubits: Qubits) -> 'Gate':
        vec = self.vec.permute(qubits)
        return Gate(vec.tensor, qubits=vec.qubits)

    @property
    def H(self) -> 'Gate':
        return Gate(self.matrix.conj().T.copy(), qubits=self.qubits)

    def asoperator(self) -> bk.BKTensor:
        """The gate as a square 2^K x 2^K device tensor (ops.py:156-158)."""
        return bk.reshape(self.tensor, [2 ** self.qubit_nb] * 2)

    # ---- action ----
    def run(self, ket: State) -> State:
        """Apply to a state: label -> axis lookup, then the gate kernel
        (ops.py:161-166).  Out of place; ValueError for an unknown qubit."""
        indices = [ket.qubits.index(q) for q in self.qubits]
        if self._host is None and self.qubit_nb > 3:
            tensor = bk.tensormul(self._dev, ket.tensor, indices)
        else:
            tensor = bk.tensormul(self.kind, ket.tensor, indices)
        return State(tensor, ket.qubits, ket.memory)

    def aschannel(self):
        """The gate's superoperator U (x) conj(U) as a Channel (ops.py:243-252)."""
        from .channels import gate_aschannel
        return gate_aschannel(self)

    def evolve(self, rho):
        """rho -> U rho U^H on a Density (ops.py:168-172): the gate kernels
        on the density's row axes (U) and column axes (conj U)."""
        from .channels import gate_evolve
        return gate_evolve(self, rho)

    def __pow__(self, t: float) -> 'Gate':
        """Fractional matrix power (ops.py:175-184)."""
        N = self.qubit_nb
        m = matpow(self.matrix, t)
        return Gate(np.reshape(m, [2] * (2 * N)), self.qubits)

    def __matmul__(self, other: 'Gate') -> 'Gate':
        """Compose: apply this gate to ``other`` (whose qubits must contain
        this gate's), computed on the GPU with the gate kernels treating the
        other gate's tensor as a 2K'-qubit vector (ops.py:187-198)."""
        if not isinstance(other, Gate):
            raise NotImplementedError()
        indices = [other.qubits.index(q) for q in self.qubits]
        tensor = bk.tensormul(self.kind, other.tensor, indices)
        return Gate(tensor=tensor, qubits=other.qubits)

    def quil(self) -> str:
        if self.name == 'Gate':
            return object.__repr__(self)
        fqubits = " " + " ".join([str(q) for q in self.qubits])
        if self.params:
            def fmt(p):
                if isinstance(p, float):
                    return ('%.10g' % p)
                return str(p)
            fparams = "(" + ", ".join(fmt(p) for p in self.params.values()) + ")"
        else:
            fparams = ""
        return "{}{}{}".format(self.name, fparams, fqubits)

    def asgate(self) -> 'Gate':
        return self

    def su(self) -> 'Gate':
        """Rescale to the special unitary group (ops.py:254-259)."""
        rank = 2 ** self.qubit_nb
        U = self.matrix.copy()
        U /= np.linalg.det(U) ** (1 / rank)
        return Gate(tensor=U, qubits=self.qubits)

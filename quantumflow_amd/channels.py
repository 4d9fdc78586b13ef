"""Density matrices and Kraus channels on the GPU.

Mirrors the reference's Density (quantumflow/states.py:264-344), Channel
(quantumflow/ops.py:265-412) and channels (quantumflow/channels.py:36-217).  A Density of N qubits holds its
matrix as a [2]*2N complex128 device tensor: axes 0..N-1 are the row (ket)
indices, N..2N-1 the column (bra) indices.  A K-qubit operator U acts as
rho -> U rho U^H, which is U on the row axes and conj(U) on the column
axes.  Where a gate kernel takes U (x) conj(U) as one 2K-qubit gate on both
sets of axes (the superoperator; Gate.aschannel, quantumflow/ops.py:
243-252), that is one pass of the state-vector gate kernels over the
2N-axis tensor; otherwise U and conj(U) are two passes.  1-qubit Kraus
channels and Channel objects run their superoperator the same way.  Outer
products, identity, diagonal, trace, partial trace and the weighted sums
of wider Kraus channels run in csrc/density.hip.
"""

from typing import Sequence, Union

import numpy as np
import torch

from . import _native
from . import backend as bk
from .ops import Gate, Operation
from .qubits import QubitVector, Qubits, qubits_count_tuple
from .states import State

__all__ = ['Density', 'mixed_density', 'random_density', 'Channel', 'Kraus',
           'UnitaryMixture', 'Depolarizing', 'Damping', 'Dephasing', 'kraus_iscomplete',
           'join_channels', 'channel_to_kraus']

MAX_QUBITS = 15        # QFX_DENSITY_MAX_QUBITS


def _ptr(t):
    return _native.ctypes.c_void_p(t.data_ptr())


def _check_n(n):
    if not 0 <= n <= MAX_QUBITS:
        raise ValueError('density matrices hold 0..%d qubits' % MAX_QUBITS)


class Density(State):
    """A density matrix of a mixed state (states.py:264-320).  ``tensor``:
    the 2^N x 2^N matrix in any shape, host or device."""

    def __init__(self, tensor, qubits: Qubits = None, memory=None) -> None:
        if qubits is None:
            tensor = bk.astensorproduct(tensor)
            qubits = range(bk.rank(tensor) // 2)
        super().__init__(tensor, qubits, memory)
        _check_n(self.qubit_nb)

    def _mat(self):
        return bk.as_device(self.tensor)

    def trace(self):
        """sum_i rho_ii as a 0-dim complex128 device tensor (states.py:277-279)."""
        out = torch.empty(1, dtype=torch.complex128, device=bk.DEVICE)
        _native.call('qfx_density_trace', bk.stream_handle(), _ptr(out),
                     _ptr(self._mat()), self.qubit_nb)
        return out[0]

    def normalize(self) -> 'Density':
        """rho / trace(rho) (states.py:296-299)."""
        tr = complex(self.trace().item())
        out = torch.zeros_like(self._mat())
        _native.call('qfx_density_axpy', bk.stream_handle(), _ptr(out),
                     _ptr(self._mat()), 1.0 / tr.real, out.numel())
        return Density(out.view([2] * (2 * self.qubit_nb)), self.qubits, self._memory)

    def probabilities(self):
        """Re rho_ii as a [2]*N float64 device tensor (states.py:302-305)."""
        n = self.qubit_nb
        out = torch.empty(2 ** n, dtype=torch.float64, device=bk.DEVICE)
        _native.call('qfx_density_diag', bk.stream_handle(), _ptr(out),
                     _ptr(self._mat()), n)
        return out.view([2] * n)

    def asoperator(self):
        """The 2^N x 2^N matrix (a device view) (states.py:307-309)."""
        return self._mat().view(2 ** self.qubit_nb, 2 ** self.qubit_nb)

    def asdensity(self) -> 'Density':
        return self

    def partial_trace(self, qubits: Qubits) -> 'Density':
        """Trace out ``qubits``; the others keep their order (states.py:
        281-285, qubits.py:201-229).  ValueError for a qubit not held or for
        tracing out every qubit."""
        new_qubits = list(self.qubits)
        for q in qubits:
            new_qubits.remove(q)
        if not new_qubits:
            raise ValueError('Cannot remove all qubits with partial_trace.')
        n = self.qubit_nb
        keep = 0
        for q in new_qubits:
            keep |= 1 << (n - 1 - self.qubits.index(q))
        m = len(new_qubits)
        out = torch.empty(4 ** m, dtype=torch.complex128, device=bk.DEVICE)
        _native.call('qfx_density_partial_trace', bk.stream_handle(), _ptr(out),
                     _ptr(self._mat()), n, keep)
        return Density(out.view([2] * (2 * m)), new_qubits, self._memory)

    def relabel(self, qubits: Qubits) -> 'Density':
        return Density(self.tensor, qubits, self._memory)

    def permute(self, qubits: Qubits) -> 'Density':
        """The same density with its axes reordered to ``qubits`` (row and
        column axes alike), still a Density (states.py:287-290)."""
        vec = self.vec.permute(qubits)
        return Density(vec.tensor, vec.qubits, self._memory)

    def update(self, memory) -> 'Density':
        mem = self.memory
        mem.update(memory)
        return Density(self.tensor, self.qubits, mem)


def mixed_density(qubits: Union[int, Qubits]) -> Density:
    """I / 2^N, filled on the device (states.py:322-326)."""
    N, qubits = qubits_count_tuple(qubits)
    _check_n(N)
    rho = torch.empty(4 ** N, dtype=torch.complex128, device=bk.DEVICE)
    _native.call('qfx_density_identity', bk.stream_handle(), _ptr(rho), N, 2.0 ** -N)
    return Density(rho.view([2] * (2 * N)), qubits)


def random_density(qubits: Union[int, Qubits]) -> Density:
    """A Hilbert-Schmidt random density (states.py:329-344): G G^H / tr for a
    Ginibre matrix G drawn on the host (numpy's global generator)."""
    N, qubits = qubits_count_tuple(qubits)
    size = (2 ** N, 2 ** N)
    g = (np.random.normal(size=size) + 1j * np.random.normal(size=size)) / np.sqrt(2.0)
    m = g @ g.conj().T
    m /= np.trace(m)
    return Density(m.reshape([2] * (2 * N)), qubits)


def state_asdensity(ket: State) -> Density:
    """|psi><psi| from a state, on the device (states.py:162-165)."""
    n = ket.qubit_nb
    _check_n(n)
    psi = bk.as_device(ket.tensor)
    rho = torch.empty(4 ** n, dtype=torch.complex128, device=bk.DEVICE)
    _native.call('qfx_density_outer', bk.stream_handle(), _ptr(rho), _ptr(psi), n)
    return Density(rho.view([2] * (2 * n)), ket.qubits, ket._memory)


def _one_pass_kind(m):
    """The kernel class of U (x) conj(U) if one gate kernel takes it, else
    None."""
    k = int(round(np.log2(m.shape[0])))
    if 2 * k > 5:
        return None
    kind = bk.classify(np.kron(m, np.conj(m)))
    return kind if 2 * k <= 3 or kind.kind != kind.DENSE else None


def evolve_passes(gate: Gate) -> int:
    """HBM passes over the density that gate_evolve makes for this gate."""
    return 1 if _one_pass_kind(gate.matrix) is not None else 2


def gate_evolve(gate: Gate, rho: Density) -> Density:
    """U rho U^H (ops.py:168-172): U on the row axes of the gate's qubits,
    conj(U) on their column axes.  When U (x) conj(U), as one 2K-qubit gate
    on both sets of axes, fits a single gate kernel (dense with 2K <= 3,
    diagonal or monomial with 2K <= 5: every 1-qubit gate, CNOT, CZ, SWAP,
    ISWAP, ...) the map is one pass over the density; otherwise two."""
    N = rho.qubit_nb
    idx = [rho.qubits.index(q) for q in gate.qubits]
    k = len(idx)
    m = gate.matrix
    cols = [N + i for i in idx]
    kind = _one_pass_kind(m)
    if kind is not None:
        t = bk.tensormul(kind, rho.tensor, idx + cols)
        return Density(t, rho.qubits, rho._memory)
    t = bk.tensormul(m.reshape([2] * (2 * k)), rho.tensor, idx)
    t = bk.tensormul(np.conj(m).reshape([2] * (2 * k)), t, cols)
    return Density(t, rho.qubits, rho._memory)


class Channel(Operation):
    """A channel as its superoperator (ops.py:265-412): a [2]*4N tensor with
    axes (rows out, columns out, rows in, columns in), i.e. the 4^N x 4^N
    matrix that maps a vectorised density (rows then columns).  For a gate
    it is U (x) conj(U) (Gate.aschannel, ops.py:243-252).  ``evolve`` and
    ``@`` run it as one 2N-qubit gate on the gate kernels."""

    def __init__(self, tensor, qubits, params=None, name=None) -> None:
        _, qubits = qubits_count_tuple(qubits)
        self._qubits = tuple(qubits)
        N = len(self._qubits)
        if isinstance(tensor, torch.Tensor):
            tensor = tensor.detach().cpu().numpy()
        m = np.asarray(tensor, dtype=np.complex128)
        if m.size != 16 ** N:
            raise ValueError('Incompatibility between tensor and qubits')
        self._mat = m.reshape(4 ** N, 4 ** N)
        self.params = params
        self._name = name if name is not None else self.__class__.__name__

    @property
    def name(self) -> str:
        return self._name

    @property
    def qubits(self) -> Qubits:
        return self._qubits

    @property
    def qubit_nb(self) -> int:
        return len(self._qubits)

    @property
    def tensor(self):
        """The superoperator as a [2]*4N device tensor (ops.py:285-288)."""
        return bk.astensorproduct(self._mat.reshape([2] * (4 * self.qubit_nb)))

    def asoperator(self) -> np.ndarray:
        """The 4^N x 4^N superoperator matrix (host)."""
        return self._mat

    def relabel(self, qubits: Qubits) -> 'Channel':
        return Channel(self._mat, qubits, self.params, self._name)

    @property
    def vec(self) -> QubitVector:
        """The superoperator as a rank-4 QubitVector on the device
        (ops.py:272)."""
        return QubitVector(self.tensor, self.qubits, rank=4)

    def permute(self, qubits: Qubits) -> 'Channel':
        """The channel with its axes reordered to ``qubits`` in all four
        super-indices (ops.py:303-306)."""
        qubits = tuple(qubits)
        N = self.qubit_nb
        idx = [self.qubits.index(q) for q in qubits]
        perm = [r * N + i for r in range(4) for i in idx]
        t = self._mat.reshape([2] * (4 * N)).transpose(perm)
        return Channel(t.reshape(4 ** N, 4 ** N), qubits)

    @property
    def H(self) -> 'Channel':
        """Conjugate transpose of the superoperator (ops.py:308-310)."""
        return Channel(self._mat.conj().T, self.qubits)

    def chi(self) -> np.ndarray:
        """The chi (process) matrix as the reference defines it: the sharp
        transpose flattened to a 4^N x 4^N matrix, i.e. equal to the Choi
        matrix (ops.py:344-349)."""
        return self.choi()

    def trace(self) -> complex:
        """Trace of the 4^N x 4^N superoperator matrix (ops.py:400-403,
        qubits.py:185-197)."""
        return complex(np.trace(self._mat))

    def partial_trace(self, qubits: Qubits) -> 'Channel':
        """Sum over the traced qubits with all four of their super-index
        axes tied to one index (qubits.py:201-227, einsum with the traced
        qubit's subscript repeated in every super-index); the kept axes stay
        in their order."""
        new_qubits = list(self.qubits)
        for q in qubits:
            new_qubits.remove(q)
        if not new_qubits:
            raise ValueError('Cannot remove all qubits with partial_trace.')
        N = self.qubit_nb
        traced = {self.qubits.index(q) for q in qubits}
        letters = list(range(4 * N))
        for i in traced:
            for r in range(1, 4):
                letters[r * N + i] = i
        out = [r * N + i for r in range(4) for i in range(N) if i not in traced]
        t = np.einsum(self._mat.reshape([2] * (4 * N)), letters, out)
        M = len(new_qubits)
        return Channel(t.reshape(4 ** M, 4 ** M), new_qubits)

    @property
    def sharp(self) -> 'Channel':
        """Swap the two bra super-indices (ops.py:313-335)."""
        N = self.qubit_nb
        t = self._mat.reshape([2 ** N] * 4).transpose(0, 2, 1, 3)
        return Channel(t.reshape(4 ** N, 4 ** N), self.qubits)

    def choi(self) -> np.ndarray:
        """The Choi matrix (ops.py:337-342)."""
        return self.sharp._mat.reshape([4 ** self.qubit_nb] * 2)

    def run(self, ket: State) -> State:
        raise TypeError()  # not possible in general (ops.py:351-352)

    def asgate(self) -> Gate:
        raise TypeError()  # not possible in general (ops.py:365-366)

    def aschannel(self) -> 'Channel':
        return self

    def evolve(self, rho: Density) -> Density:
        """The superoperator on the channel's row and column axes of the
        density, one gate-kernel pass (ops.py:354-363)."""
        N = rho.qubit_nb
        idx = [rho.qubits.index(q) for q in self.qubits]
        kind = bk.classify(self._mat)
        t = bk.tensormul(kind, rho.tensor, idx + [N + i for i in idx])
        return Density(t, rho.qubits, rho._memory)

    def __add__(self, other) -> 'Channel':
        if isinstance(other, Channel):
            if not self.qubits == other.qubits:
                raise ValueError("Qubits must be identical")
            return Channel(self._mat + other._mat, self.qubits)
        raise NotImplementedError()

    def __mul__(self, other) -> 'Channel':
        return Channel(self._mat * other, self.qubits)

    def __matmul__(self, other: 'Channel') -> 'Channel':
        """This channel after ``other`` (whose qubits hold this one's): this
        superoperator on the output axes of the other's tensor, as a gate on
        a 4N-axis state (ops.py:384-398)."""
        if not isinstance(other, Channel):
            raise NotImplementedError()
        N = other.qubit_nb
        idx = [other.qubits.index(q) for q in self.qubits]
        t = bk.tensormul(bk.classify(self._mat), other.tensor, idx + [N + i for i in idx])
        return Channel(t, other.qubits)


def gate_aschannel(gate: Gate) -> Channel:
    """U (x) conj(U) in (row, column) order (ops.py:243-252)."""
    m = gate.matrix
    return Channel(np.kron(m, np.conj(m)), gate.qubits)


class Kraus(Operation):
    """A Kraus representation of a channel (channels.py:36-105):
    rho -> sum_k w_k K_k rho K_k^H."""

    def __init__(self, operators: Sequence[Gate], weights: Sequence[float] = None) -> None:
        self.operators = list(operators)
        if weights is None:
            weights = [1.0] * len(self.operators)
        self.weights = tuple(float(w) for w in weights)

    @property
    def qubits(self) -> Qubits:
        return tuple(sorted(set(q for op in self.operators for q in op.qubits)))

    def asgate(self) -> Gate:
        raise TypeError('Not possible in general')

    def aschannel(self) -> Channel:
        """The channel's superoperator (channels.py:52-62)."""
        return Channel(self.superoperator(), self.qubits)

    def superoperator(self):
        """sum_k w_k K_k (x) conj(K_k) on the channel's qubits (rows then
        columns): the channel's action on vectorised densities, as a
        4^K x 4^K host matrix (the map of Kraus.aschannel, channels.py:
        52-62)."""
        qubits = self.qubits
        K = len(qubits)
        S = np.zeros((4 ** K, 4 ** K), dtype=np.complex128)
        for op, w in zip(self.operators, self.weights):
            m = _embed(op.matrix, [qubits.index(q) for q in op.qubits], K)
            S += w * np.kron(m, np.conj(m))
        return S

    def evolve(self, rho: Density) -> Density:
        """Weighted sum of the operators' evolutions (channels.py:74-80).
        When one gate kernel takes the channel's superoperator (1-qubit
        channels: a dense 2-qubit gate on the row and column axes), that is
        one pass over the density; otherwise each operator's evolution is
        accumulated by weighted sums."""
        qubits = self.qubits
        if 2 * len(qubits) <= 5:
            kind = bk.classify(self.superoperator())
            if 2 * len(qubits) <= 3 or kind.kind != kind.DENSE:
                N = rho.qubit_nb
                idx = [rho.qubits.index(q) for q in qubits]
                t = bk.tensormul(kind, rho.tensor, idx + [N + i for i in idx])
                return Density(t, rho.qubits, rho._memory)
        out = torch.zeros(4 ** rho.qubit_nb, dtype=torch.complex128, device=bk.DEVICE)
        for op, w in zip(self.operators, self.weights):
            term = bk.as_device(op.evolve(rho).tensor)
            _native.call('qfx_density_axpy', bk.stream_handle(), _ptr(out), _ptr(term),
                         w, out.numel())
        return Density(out.view([2] * (2 * rho.qubit_nb)), rho.qubits, rho._memory)

    def run(self, ket: State) -> State:
        """One Kraus branch drawn with probability w_k |K_k psi|^2, the state
        renormalised (channels.py:64-71)."""
        res = [op.run(ket) for op in self.operators]
        probs = np.array([float(np.abs(bk.evaluate(r.norm()))) * w
                          for r, w in zip(res, self.weights)])
        probs /= np.sum(probs)
        pick = res[np.random.choice(len(res), p=probs)]
        return pick.normalize()

    @property
    def H(self) -> 'Kraus':
        return Kraus([op.H for op in self.operators], self.weights)


class UnitaryMixture(Kraus):
    """A convex mixture of unitaries (channels.py:108-125)."""

    def asgate(self) -> Gate:
        return self.operators[np.random.choice(len(self.operators), p=self.weights)]

    def run(self, ket: State) -> State:
        return self.asgate().run(ket)


class Depolarizing(UnitaryMixture):
    """I, X, Y, Z with weights 1-p, p/3, p/3, p/3 (channels.py:128-138)."""

    def __init__(self, prob: float, q0) -> None:
        from .stdgates import I, X, Y, Z
        super().__init__([I(q0), X(q0), Y(q0), Z(q0)],
                         [1 - prob, prob / 3.0, prob / 3.0, prob / 3.0])


class Damping(Kraus):
    """Amplitude damping: [[1, 0], [0, sqrt(1-p)]] and [[0, sqrt(p)], [0, 0]]
    (channels.py:141-153)."""

    def __init__(self, prob: float, q0) -> None:
        k0 = Gate(np.array([[1.0, 0.0], [0.0, np.sqrt(1 - prob)]]), qubits=[q0])
        k1 = Gate(np.array([[0.0, np.sqrt(prob)], [0.0, 0.0]]), qubits=[q0])
        super().__init__([k0, k1])


class Dephasing(UnitaryMixture):
    """I and Z with weights 1-p/2, p/2 (channels.py:156-165)."""

    def __init__(self, prob: float, q0) -> None:
        from .stdgates import I, Z
        super().__init__([I(q0), Z(q0)], [1 - prob / 2, prob / 2])


def _embed(mat, pos, N):
    """A 2^k x 2^k operator on the qubits at positions ``pos`` (position 0 =
    most significant) as a 2^N x 2^N matrix (identity elsewhere)."""
    k = len(pos)
    others = [i for i in range(N) if i not in pos]
    t = np.multiply.outer(mat.reshape([2] * (2 * k)),
                          np.eye(2 ** (N - k)).reshape([2] * (2 * (N - k))))
    # axes of t: operator out (k), operator in (k), identity out, identity in
    out_ax = {p: j for j, p in enumerate(pos)}
    out_ax.update({o: 2 * k + j for j, o in enumerate(others)})
    in_ax = {p: k + j for j, p in enumerate(pos)}
    in_ax.update({o: 2 * k + (N - k) + j for j, o in enumerate(others)})
    perm = [out_ax[i] for i in range(N)] + [in_ax[i] for i in range(N)]
    return np.transpose(t, perm).reshape(2 ** N, 2 ** N)


def kraus_iscomplete(kraus: Kraus) -> bool:
    """sum_k w_k K_k^H K_k == I on the channel's qubits (channels.py:200-217),
    from the operators' host matrices."""
    qubits = kraus.qubits
    N = len(qubits)
    total = np.zeros((2 ** N, 2 ** N), dtype=np.complex128)
    for op, w in zip(kraus.operators, kraus.weights):
        kk = op.matrix.conj().T @ op.matrix
        total += w * _embed(kk, [qubits.index(q) for q in op.qubits], N)
    return bool(np.allclose(total, np.eye(2 ** N)))


def join_channels(*channels: Channel) -> Channel:
    """One channel on the union of the channels' (disjoint) qubits, in the
    given order (channels.py:170-176): the tensor product of the
    superoperators with the row and column super-indices regrouped."""
    mats = [c.asoperator() for c in channels]
    ns = [c.qubit_nb for c in channels]
    qubits = [q for c in channels for q in c.qubits]
    S = np.ones((1, 1), dtype=np.complex128)
    N = 0
    for m, k in zip(mats, ns):
        # S: (ro, co, ri, ci) over N qubits; m: the same over k qubits
        a = S.reshape([2 ** N] * 4)
        b = m.reshape([2 ** k] * 4)
        t = np.einsum('abcd,efgh->aebfcgdh', a, b)
        N += k
        S = t.reshape(4 ** N, 4 ** N)
    return Channel(S, qubits)


def channel_to_kraus(chan: Channel) -> Kraus:
    """Kraus operators of a channel from the eigenvectors of its Choi matrix
    (channels.py:179-205): K_i = sqrt(l_i) * eigenvector i as a 2^N x 2^N
    matrix, for the non-zero eigenvalues l_i (a CP map has them >= 0)."""
    N = chan.qubit_nb
    choi = chan.choi()
    evals, evecs = np.linalg.eigh((choi + choi.conj().T) / 2)
    if np.any(evals < -1e-10):
        raise ValueError('channel is not completely positive')
    ops = []
    for i in range(len(evals)):
        if not np.isclose(evals[i], 0.0):
            mat = evecs[:, i].reshape(2 ** N, 2 ** N) * np.sqrt(evals[i])
            ops.append(Gate(mat, qubits=chan.qubits))
    return Kraus(ops)

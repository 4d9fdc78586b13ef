"""Standard gates (quantumflow/stdgates.py:1-1001).

Every class builds its 2^K x 2^K matrix on the host from the closed form of
the reference class named in its docstring, with the same numpy expressions
(so matrices agree bit for bit).  Constructor signatures follow the
reference: parameters first, then qubits (default 0, 1, 2).
"""

import copy
from math import pi, sqrt

import numpy as np

from .gates import I
from .ops import Gate

__all__ = ['I', 'X', 'Y', 'Z', 'H', 'S', 'T', 'PHASE', 'RX', 'RY', 'RZ', 'CZ',
           'CNOT', 'SWAP', 'ISWAP', 'CPHASE00', 'CPHASE01', 'CPHASE10',
           'CPHASE', 'PSWAP', 'CCNOT', 'CSWAP', 'RN', 'TX', 'TY', 'TZ', 'TH',
           'ZYZ', 'CAN', 'XX', 'YY', 'ZZ', 'PISWAP', 'EXCH', 'CANONICAL',
           'S_H', 'T_H', 'STDGATES']


def _cis(x):
    return np.exp(x * 1.0j)


class _StdGate(Gate):
    """Base of the standard gates: parses (params..., qubits...)."""
    _pnames = ()
    _nq = 1

    def __init__(self, *args, **kwargs):
        np_ = len(self._pnames)
        params = list(args[:np_])
        qubits = list(args[np_:])
        for i, pn in enumerate(self._pnames):
            if pn in kwargs:
                if i < len(params):
                    raise TypeError('parameter %s given twice' % pn)
                params.append(kwargs.pop(pn))
        if len(params) != np_:
            raise TypeError('%s takes %d parameters' % (type(self).__name__, np_))
        for j in range(self._nq):
            key = 'q%d' % j
            if key in kwargs:
                qubits.append(kwargs.pop(key))
        if kwargs:
            raise TypeError('unexpected arguments %s' % list(kwargs))
        if not qubits:
            qubits = list(range(self._nq))
        if len(qubits) != self._nq:
            raise TypeError('%s acts on %d qubits' % (type(self).__name__, self._nq))
        pdict = dict(zip(self._pnames, self._norm_params(params)))
        super().__init__(self._matrix(*pdict.values()), qubits,
                         pdict if pdict else None)

    def _norm_params(self, params):
        return params

    def _rebuild(self, *params):
        return type(self)(*params, *self.qubits)


class _Hermitian(_StdGate):
    @property
    def H(self) -> Gate:
        return copy.copy(self)


class _NegH(_StdGate):
    """H negates every parameter; t-th power scales them."""

    @property
    def H(self) -> Gate:
        return self._rebuild(*[-p for p in self.params.values()])

    def __pow__(self, t):
        return self._rebuild(*[p * t for p in self.params.values()])


# ---- 1-qubit -----------------------------------------------------------------

class X(_Hermitian):
    """Pauli-X (stdgates.py:32-48)."""
    def _matrix(self):
        return [[0, 1], [1, 0]]

    def __pow__(self, t):
        return TX(t, *self.qubits)


class Y(_Hermitian):
    """Pauli-Y (stdgates.py:51-69)."""
    def _matrix(self):
        return np.asarray([[0, -1.0j], [1.0j, 0]])

    def __pow__(self, t):
        return TY(t, *self.qubits)


class Z(_Hermitian):
    """Pauli-Z (stdgates.py:72-88)."""
    def _matrix(self):
        return [[1, 0], [0, -1.0]]

    def __pow__(self, t):
        return TZ(t, *self.qubits)


class H(_Hermitian):
    """Hadamard (stdgates.py:91-109)."""
    def _matrix(self):
        return np.asarray([[1, 1], [1, -1]]) / sqrt(2)

    def __pow__(self, t):
        return TH(t, *self.qubits)


class S(_StdGate):
    """S = PHASE(pi/2) (stdgates.py:112-130)."""
    def _matrix(self):
        return np.asarray([[1.0, 0.0], [0.0, 1.0j]])

    @property
    def H(self):
        return S_H(*self.qubits)

    def __pow__(self, t):
        return PHASE(pi / 2 * t, *self.qubits)


class T(_StdGate):
    """T = PHASE(pi/4) (stdgates.py:133-151)."""
    def _matrix(self):
        return [[1.0, 0.0], [0.0, _cis(pi / 4.0)]]

    @property
    def H(self):
        return T_H(*self.qubits)

    def __pow__(self, t):
        return PHASE(pi / 4 * t, *self.qubits)


class S_H(_StdGate):
    """Inverse of S (stdgates.py:566-582)."""
    def _matrix(self):
        return np.asarray([[1.0, 0.0], [0.0, -1.0j]])

    @property
    def H(self):
        return S(*self.qubits)

    def __pow__(self, t):
        return PHASE(-pi / 2 * t, *self.qubits)


class T_H(_StdGate):
    """Inverse of T (stdgates.py:585-601)."""
    def _matrix(self):
        return [[1.0, 0.0], [0.0, _cis(-pi / 4.0)]]

    @property
    def H(self):
        return T(*self.qubits)

    def __pow__(self, t):
        return PHASE(-pi / 4 * t, *self.qubits)


class PHASE(_StdGate):
    """diag(1, e^{i theta}) (stdgates.py:154-176)."""
    _pnames = ('theta',)

    def _matrix(self, theta):
        return [[1.0, 0.0], [0.0, _cis(theta)]]

    @property
    def H(self):
        theta = self.params['theta']
        return PHASE(2. * pi - theta % (2. * pi), *self.qubits)

    def __pow__(self, t):
        return PHASE(self.params['theta'] * t, *self.qubits)


class RX(_NegH):
    """exp(-i theta X / 2) (stdgates.py:179-205)."""
    _pnames = ('theta',)

    def _matrix(self, theta):
        return [[np.cos(theta / 2), -1.0j * np.sin(theta / 2)],
                [-1.0j * np.sin(theta / 2), np.cos(theta / 2)]]


class RY(_NegH):
    """exp(-i theta Y / 2) (stdgates.py:208-233)."""
    _pnames = ('theta',)

    def _matrix(self, theta):
        return [[np.cos(theta / 2.0), -np.sin(theta / 2.0)],
                [np.sin(theta / 2.0), np.cos(theta / 2.0)]]


class RZ(_NegH):
    """exp(-i theta Z / 2) (stdgates.py:236-262)."""
    _pnames = ('theta',)

    def _matrix(self, theta):
        return [[np.exp(-theta * 0.5j), 0], [0, np.exp(theta * 0.5j)]]


class RN(_StdGate):
    """Rotation by theta about axis (nx, ny, nz) (stdgates.py:604-636)."""
    _pnames = ('theta', 'nx', 'ny', 'nz')

    def _matrix(self, theta, nx, ny, nz):
        c, s = np.cos(theta / 2), np.sin(theta / 2)
        return [[c - 1j * s * nz, -1j * s * nx - s * ny],
                [-1j * s * nx + s * ny, c + 1j * s * nz]]

    @property
    def H(self):
        theta, nx, ny, nz = self.params.values()
        return RN(-theta, nx, ny, nz, *self.qubits)

    def __pow__(self, t):
        theta, nx, ny, nz = self.params.values()
        return RN(t * theta, nx, ny, nz, *self.qubits)


class _TPow(_NegH):
    """Powers of a Pauli: parameter t reduced mod 2 (stdgates.py:685,717,748)."""
    _pnames = ('t',)

    def _norm_params(self, params):
        return [params[0] % 2]


class TX(_TPow):
    """X^t = e^{i pi t/2} RX(pi t) (stdgates.py:639-669)."""
    def _matrix(self, t):
        ct = pi * t
        ph = np.exp(0.5j * ct)
        return [[ph * np.cos(ct / 2), ph * -1.0j * np.sin(ct / 2)],
                [ph * -1.0j * np.sin(ct / 2), ph * np.cos(ct / 2)]]


class TY(_TPow):
    """Y^t (stdgates.py:672-704)."""
    def _matrix(self, t):
        ct = pi * t
        ph = np.exp(0.5j * ct)
        return [[ph * np.cos(ct / 2.0), ph * -np.sin(ct / 2.0)],
                [ph * np.sin(ct / 2.0), ph * np.cos(ct / 2.0)]]


class TZ(_TPow):
    """Z^t (stdgates.py:707-736)."""
    def _matrix(self, t):
        ct = pi * t
        ph = np.exp(0.5j * ct)
        return [[ph * np.exp(-ct * 0.5j), 0], [0, ph * np.exp(ct * 0.5j)]]


class TH(_NegH):
    """H^t (stdgates.py:761-793)."""
    _pnames = ('t',)

    def _matrix(self, t):
        th = pi * t
        ph = np.exp(0.5j * th)
        a = (ph * 1.0j * np.sin(th / 2)) / sqrt(2)
        return [[ph * np.cos(th / 2) - a, -a], [-a, ph * np.cos(th / 2) + a]]


class ZYZ(_StdGate):
    """Z^t2 Y^t1 Z^t0 (stdgates.py:796-834)."""
    _pnames = ('t0', 't1', 't2')

    def _matrix(self, t0, t1, t2):
        c0, c1, c2 = pi * t0, pi * t1, pi * t2
        return [[_cis(-0.5 * c2 - 0.5 * c0) * np.cos(0.5 * c1),
                 -_cis(-0.5 * c2 + 0.5 * c0) * np.sin(0.5 * c1)],
                [_cis(0.5 * c2 - 0.5 * c0) * np.sin(0.5 * c1),
                 _cis(0.5 * c2 + 0.5 * c0) * np.cos(0.5 * c1)]]

    @property
    def H(self):
        t0, t1, t2 = self.params.values()
        return ZYZ(-t2, -t1, -t0, *self.qubits)


# ---- 2-qubit -----------------------------------------------------------------

class CZ(_Hermitian):
    """Controlled-Z (stdgates.py:267-288)."""
    _nq = 2

    def _matrix(self):
        return [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]]


class CNOT(_Hermitian):
    """Controlled-NOT, control = first qubit (stdgates.py:291-312)."""
    _nq = 2

    def _matrix(self):
        return [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]


class SWAP(_Hermitian):
    """Swap (stdgates.py:315-338)."""
    _nq = 2

    def _matrix(self):
        return [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]


class ISWAP(_StdGate):
    """iSwap (stdgates.py:341-362)."""
    _nq = 2

    def _matrix(self):
        return np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])


class _CPhaseBase(_NegH):
    _nq = 2
    _pnames = ('theta',)
    _slot = 3

    def _matrix(self, theta):
        d = [1.0, 1.0, 1.0, 1.0]
        d[self._slot] = np.exp(1j * theta)
        return np.diag(d)


class CPHASE00(_CPhaseBase):
    """diag(e^{i theta}, 1, 1, 1) (stdgates.py:365-386)."""
    _slot = 0


class CPHASE01(_CPhaseBase):
    """diag(1, e^{i theta}, 1, 1) (stdgates.py:389-410)."""
    _slot = 1


class CPHASE10(_CPhaseBase):
    """diag(1, 1, e^{i theta}, 1) (stdgates.py:413-434)."""
    _slot = 2


class CPHASE(_CPhaseBase):
    """diag(1, 1, 1, e^{i theta}) (stdgates.py:437-456)."""
    _slot = 3


class PSWAP(_StdGate):
    """Parametric swap (stdgates.py:459-482)."""
    _nq = 2
    _pnames = ('theta',)

    def _matrix(self, theta):
        e = np.exp(theta * 1.0j)
        return [[1, 0, 0, 0], [0, 0, e, 0], [0, e, 0, 0], [0, 0, 0, 1]]

    @property
    def H(self):
        theta = self.params['theta']
        return PSWAP(2. * pi - theta % (2. * pi), *self.qubits)


class PISWAP(_NegH):
    """Parametric iSwap from the XY interaction (stdgates.py:485-514)."""
    _nq = 2
    _pnames = ('theta',)

    def _matrix(self, theta):
        c, s = np.cos(2 * theta), np.sin(2 * theta) * 1j
        return [[1, 0, 0, 0], [0, c, s, 0], [0, s, c, 0], [0, 0, 0, 1]]


class XX(_NegH):
    """exp(-i pi t X.X / 2) (stdgates.py:860-887)."""
    _nq = 2
    _pnames = ('t',)

    def _matrix(self, t):
        th = pi * t
        c, s = np.cos(th / 2), np.sin(th / 2)
        return [[c, 0, 0, -1.0j * s], [0, c, -1.0j * s, 0],
                [0, -1.0j * s, c, 0], [-1.0j * s, 0, 0, c]]


class YY(_NegH):
    """exp(-i pi t Y.Y / 2) (stdgates.py:890-913)."""
    _nq = 2
    _pnames = ('t',)

    def _matrix(self, t):
        th = pi * t
        c, s = np.cos(th / 2), np.sin(th / 2)
        return [[c, 0, 0, 1.0j * s], [0, c, -1.0j * s, 0],
                [0, -1.0j * s, c, 0], [1.0j * s, 0, 0, c]]


class ZZ(_NegH):
    """exp(-i pi t Z.Z / 2) (stdgates.py:944-968)."""
    _nq = 2
    _pnames = ('t',)

    def _matrix(self, t):
        th = pi * t
        return np.diag([_cis(-th / 2), _cis(th / 2), _cis(th / 2), _cis(-th / 2)])


class CAN(_NegH):
    """Canonical gate ZZ(tz) YY(ty) XX(tx) (stdgates.py:839-857)."""
    _nq = 2
    _pnames = ('tx', 'ty', 'tz')

    def _matrix(self, tx, ty, tz):
        xx = np.asarray(XX(tx).matrix)
        yy = np.asarray(YY(ty).matrix)
        zz = np.asarray(ZZ(tz).matrix)
        return zz @ (yy @ xx)


class CANONICAL(CAN):
    """Deprecated alias of CAN (stdgates.py:852-855)."""


class EXCH(_NegH):
    """CAN(t, t, t) (stdgates.py:971-989)."""
    _nq = 2
    _pnames = ('t',)

    def _matrix(self, t):
        return CAN(t, t, t).matrix


# ---- 3-qubit -----------------------------------------------------------------

class CCNOT(_Hermitian):
    """Toffoli (stdgates.py:519-556)."""
    _nq = 3

    def _matrix(self):
        m = np.eye(8)
        m[[6, 7]] = m[[7, 6]]
        return m


class CSWAP(_Hermitian):
    """Fredkin (stdgates.py:559-563)."""
    _nq = 3

    def _matrix(self):
        m = np.eye(8)
        m[[5, 6]] = m[[6, 5]]
        return m


GATESET = frozenset([I, X, Y, Z, H, S, T, PHASE, RX, RY, RZ, CZ, CNOT, SWAP, ISWAP,
                     CPHASE00, CPHASE01, CPHASE10, CPHASE, PSWAP, CCNOT, CSWAP,
                     PISWAP, RN, TX, TY, TZ, TH, ZYZ, CAN, XX, YY, ZZ, EXCH,
                     S_H, T_H])

STDGATES = {gate_class.__name__: gate_class for gate_class in GATESET}

"""Addresses of classical memory (quantumflow/cbits.py:1-80).

A ``Register`` names a region of classical memory with a Quil data type;
indexing it gives an ``Addr``, the key under which a State's (or
Density's) ``memory`` dictionary holds a value.  Measurement outcomes,
program counters and jump targets all live there.  Both classes compare
and hash by value, so equal addresses built separately find the same cell.
"""

from functools import total_ordering
from typing import Any, Hashable

__all__ = ['Register', 'Addr']

DTYPE = {'BIT', 'REAL', 'INT', 'OCTET', 'ANY'}

DEFAULT_REGION = 'ro'


@total_ordering
class Register:
    """A named region of classical memory of one Quil type
    (cbits.py:21-48)."""

    def __init__(self, name: str = DEFAULT_REGION, dtype: str = 'BIT') -> None:
        if dtype not in DTYPE:
            raise AssertionError('unknown classical data type %r' % (dtype,))
        self.name = name
        self.dtype = dtype

    def _key(self):
        return (self.dtype, self.name)

    def __getitem__(self, key: Hashable) -> 'Addr':
        return Addr(self, key)

    def __lt__(self, other: Any) -> bool:
        if not isinstance(other, Register):
            return NotImplemented
        return self._key() < other._key()

    def __eq__(self, other: Any) -> bool:
        if not isinstance(other, Register):
            return NotImplemented
        return self._key() == other._key()

    def __hash__(self) -> int:
        return hash(self._key())

    def __repr__(self) -> str:
        return 'Register(%r, %r)' % (self.name, self.dtype)


@total_ordering
class Addr:
    """One cell of a register (cbits.py:51-80); printed as ``name[key]``."""

    def __init__(self, register: Register, key: Hashable) -> None:
        self.register = register
        self.key = key

    @property
    def dtype(self) -> str:
        return self.register.dtype

    def _key(self):
        return (self.register, self.key)

    def __str__(self) -> str:
        return '%s[%s]' % (self.register.name, self.key)

    def __repr__(self) -> str:
        return repr(self.register) + '[' + repr(self.key) + ']'

    def __lt__(self, other: Any) -> bool:
        if not isinstance(other, Addr):
            return NotImplemented
        return self._key() < other._key()

    def __eq__(self, other: Any) -> bool:
        if not isinstance(other, Addr):
            return NotImplemented
        return self._key() == other._key()

    def __hash__(self) -> int:
        return hash(self._key())

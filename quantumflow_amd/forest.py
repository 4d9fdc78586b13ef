"""Quil text in, and a QVM-style machine over the device state
(quantumflow/forest/__init__.py:49-64, 216-347, 369-444).

The reference reads Quil through pyQuil's parser and converts pyQuil's
instruction objects into this package's instructions.  pyQuil is not a
dependency here, so ``quil_to_program`` parses the text itself: a line
splitter (comments, ``;`` separators), a tokenizer per instruction and a
recursive-descent parser for parameter expressions.  It produces the
instructions the reference's conversion produces for each Quil form:

* a gate whose name is a standard gate becomes that gate on the given
  qubits; any other name becomes a ``Call``;
* memory references become ``Register(name)[offset]`` cells (a bare name is
  offset 0, the legacy ``[n]`` form is ``ro[n]``), with the register's
  default BIT type, as the reference's conversion builds them;
* ``TRUE``/``FALSE`` become ``MOVE`` of 1/0, ``OR`` becomes ``IOR``;
* ``DEFGATE``, ``DEFCIRCUIT`` and ``CONVERT`` raise ``ValueError``, as the
  reference's conversion does for instructions it does not know;
* text that is not Quil raises ``QuilParseError`` (a ``RuntimeError``).

``QuantumFlowQVM`` keeps the reference's load/run/wait/read status machine;
the program runs on the GPU state through ``Program.run``.
"""

import cmath
import math
import re
from typing import Any, List, Sequence, Tuple

from .cbits import Register
from .programs import (Call, Declare, Halt, Include, Jump, JumpUnless, JumpWhen, Label, Load,
                       Nop, Pragma, Program, Store, Wait)
from .states import State
from .stdgates import STDGATES
from .stdops import (EQ, GE, GT, LE, LT, Add, And, Div, Exchange, Ior, Measure, Move, Mul, Neg,
                     Not, Reset, Sub, Xor)

__all__ = ['QUIL_GATES', 'QUIL_RESERVED_WORDS', 'QuilParseError', 'quil_to_program',
           'QuantumFlowQVM']

# Names of Quil-compatible standard gates (forest/__init__.py:49-52)
QUIL_GATES = {'I', 'X', 'Y', 'Z', 'H', 'S', 'T', 'PHASE',
              'RX', 'RY', 'RZ', 'CZ', 'CNOT', 'SWAP',
              'ISWAP', 'CPHASE00', 'CPHASE01', 'CPHASE10',
              'CPHASE', 'PSWAP', 'CCNOT', 'CSWAP', 'PISWAP'}

# Quil keywords (forest/__init__.py:57-61)
QUIL_RESERVED_WORDS = ['DEFGATE', 'DEFCIRCUIT', 'MEASURE', 'LABEL', 'HALT',
                       'JUMP', 'JUMP-WHEN', 'JUMP-UNLESS', 'RESET', 'WAIT',
                       'NOP', 'INCLUDE', 'PRAGMA', 'DECLARE', 'NEG', 'NOT',
                       'AND', 'IOR', 'XOR', 'MOVE', 'EXCHANGE', 'CONVERT',
                       'ADD', 'SUB', 'MUL', 'DIV', 'EQ', 'GT', 'GE', 'LT',
                       'LE', 'LOAD', 'STORE', 'TRUE', 'FALSE', 'OR']

DEFAULT_REGION = 'ro'      # the legacy ``[n]`` address form names this region


class QuilParseError(RuntimeError):
    """Quil text that does not parse (``RuntimeError``, as the reference's
    parser raises)."""


# ---------------------------------------------------------------------------
# Lines and tokens

def _split_instructions(text: str) -> List[str]:
    """Source lines, less comments, split at ``;``, blank ones dropped.
    ``#`` and ``;`` inside a double-quoted string are text."""
    out = []
    for line in text.splitlines():
        cur, quoted = [], False
        for ch in line:
            if ch == '"':
                quoted = not quoted
            elif not quoted and ch == '#':
                break
            elif not quoted and ch == ';':
                out.append(''.join(cur))
                cur = []
                continue
            cur.append(ch)
        out.append(''.join(cur))
    return [s.strip() for s in out if s.strip()]


_TOKEN = re.compile(r'''
    \s*(?:
      (?P<string>"[^"]*")
    | (?P<label>@[A-Za-z_][A-Za-z0-9_\-]*)
    | (?P<addr>(?:[A-Za-z_][A-Za-z0-9_\-]*)?\[\s*\d+\s*\])
    | (?P<number>[+-]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?)
    | (?P<ident>[A-Za-z_](?:[A-Za-z0-9_\-]*[A-Za-z0-9_])?)
    | (?P<bad>\S+)
    )''', re.VERBOSE)


def _tokens(text: str) -> List[Tuple[str, str]]:
    toks = []
    for m in _TOKEN.finditer(text):
        kind = m.lastgroup
        if kind is None:
            continue
        toks.append((kind, m.group(kind)))
    return toks


def _number(text: str):
    try:
        return int(text)
    except ValueError:
        return float(text)


def _addr(tok: Tuple[str, str]):
    """A memory reference as the reference's conversion builds it:
    Register(name)[offset] (forest/__init__.py:224-227)."""
    kind, text = tok
    if kind == 'addr':
        name, idx = text.split('[')
        return Register(name or DEFAULT_REGION)[int(idx.rstrip(']').strip())]
    if kind == 'ident' and text not in QUIL_RESERVED_WORDS:
        return Register(text)[0]
    raise QuilParseError('expected a memory reference, got {!r}'.format(text))


def _operand(tok: Tuple[str, str]):
    """A memory reference or a numeric literal."""
    if tok[0] == 'number':
        return _number(tok[1])
    return _addr(tok)


def _qubit(tok: Tuple[str, str]) -> int:
    kind, text = tok
    if kind == 'number' and re.fullmatch(r'\d+', text):
        return int(text)
    raise QuilParseError('expected a qubit index, got {!r}'.format(text))


def _label(tok: Tuple[str, str]) -> str:
    if tok[0] != 'label':
        raise QuilParseError('expected a @label, got {!r}'.format(tok[1]))
    return tok[1][1:]


# ---------------------------------------------------------------------------
# Parameter expressions: numbers, pi, i, + - * / ^, unary signs, and the
# functions SIN COS SQRT EXP CIS.  Following Quil's grammar, a sign binds
# tighter than ^ (so -2^2 is 4), ^ is right-associative and binds tighter
# than * and /, which bind tighter than + and -.

_EXPR_TOKEN = re.compile(r'\s*(?:(\d+\.\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|\d+(?:[eE][+-]?\d+)?)'
                         r'|([A-Za-z_][A-Za-z0-9_]*)|(%[A-Za-z_][A-Za-z0-9_]*)|([-+*/^(),]))')


def _sqrt(x):
    if isinstance(x, complex) or x < 0:
        return cmath.sqrt(x)
    return math.sqrt(x)


def _real_or_complex(fr, fc):
    return lambda x: fc(x) if isinstance(x, complex) else fr(x)


_FUNCS = {'SIN': _real_or_complex(math.sin, cmath.sin),
          'COS': _real_or_complex(math.cos, cmath.cos),
          'EXP': _real_or_complex(math.exp, cmath.exp),
          'SQRT': _sqrt,
          'CIS': lambda x: cmath.exp(1j * x)}


class _Expr:
    def __init__(self, text: str) -> None:
        self.toks = []
        pos = 0
        text = text.rstrip()
        while pos < len(text):
            m = _EXPR_TOKEN.match(text, pos)
            if m is None or m.end() == pos:
                raise QuilParseError('cannot parse parameter {!r}'.format(text))
            num, name, var, op = m.groups()
            if num is not None:
                self.toks.append(('num', _number(num)))
            elif name is not None:
                self.toks.append(('name', name))
            elif var is not None:
                raise QuilParseError('parametric variables ({}) are not supported'.format(var))
            else:
                self.toks.append(('op', op))
            pos = m.end()
        self.i = 0

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def take(self):
        tok = self.peek()
        self.i += 1
        return tok

    def expect(self, op):
        if self.take() != ('op', op):
            raise QuilParseError('expected {!r} in a parameter expression'.format(op))

    def parse_list(self) -> List[Any]:
        vals = [self.sum()]
        while self.peek() == ('op', ','):
            self.take()
            vals.append(self.sum())
        if self.i != len(self.toks):
            raise QuilParseError('trailing text in a parameter expression')
        return vals

    def sum(self):
        v = self.product()
        while self.peek() in (('op', '+'), ('op', '-')):
            op = self.take()[1]
            w = self.product()
            v = v + w if op == '+' else v - w
        return v

    def product(self):
        v = self.power()
        while self.peek() in (('op', '*'), ('op', '/')):
            op = self.take()[1]
            w = self.power()
            v = v * w if op == '*' else v / w
        return v

    def power(self):
        v = self.signed()
        if self.peek() == ('op', '^'):
            self.take()
            v = v ** self.power()
        return v

    def signed(self):
        tok = self.peek()
        if tok in (('op', '+'), ('op', '-')):
            self.take()
            v = self.signed()
            return +v if tok[1] == '+' else -v
        return self.atom()

    def atom(self):
        kind, val = self.take()
        if kind == 'num':
            if self.peek() == ('name', 'i'):          # imaginary literal: 2.3 i
                self.take()
                return complex(0, val)
            return val
        if kind == 'name':
            if val in ('pi', 'PI'):
                return math.pi
            if val == 'i':
                return 1j
            fn = _FUNCS.get(val.upper())
            if fn is not None:
                self.expect('(')
                arg = self.sum()
                self.expect(')')
                return fn(arg)
            raise QuilParseError('unknown name {!r} in a parameter expression'.format(val))
        if (kind, val) == ('op', '('):
            v = self.sum()
            self.expect(')')
            return v
        raise QuilParseError('unexpected {!r} in a parameter expression'.format(val))


def _split_params(text: str):
    """``NAME(params) rest`` -> (NAME, [values], rest); ``NAME rest`` ->
    (NAME, [], rest)."""
    m = re.match(r'([A-Za-z_](?:[A-Za-z0-9_\-]*[A-Za-z0-9_])?)\s*(.*)$', text, re.S)
    if m is None:
        raise QuilParseError('cannot parse {!r}'.format(text))
    name, rest = m.group(1), m.group(2)
    if not rest.startswith('('):
        return name, [], rest
    depth = 0
    for k, ch in enumerate(rest):
        depth += ch == '('
        depth -= ch == ')'
        if depth == 0:
            return name, _Expr(rest[1:k]).parse_list(), rest[k + 1:]
    raise QuilParseError('unbalanced parentheses in {!r}'.format(text))


# ---------------------------------------------------------------------------
# Instructions

_BINARY = {'AND': And, 'IOR': Ior, 'OR': Ior, 'XOR': Xor, 'MOVE': Move, 'EXCHANGE': Exchange,
           'ADD': Add, 'SUB': Sub, 'MUL': Mul, 'DIV': Div}
_COMPARE = {'EQ': EQ, 'GT': GT, 'GE': GE, 'LT': LT, 'LE': LE}
_NULLARY = {'HALT': Halt, 'WAIT': Wait, 'NOP': Nop}


def _nargs(head: str, args, n: int) -> None:
    if len(args) != n:
        raise QuilParseError('{} takes {} operand(s), got {}'.format(head, n, len(args)))


def _instruction(text: str):
    head_tok = _tokens(text)[0]
    head = head_tok[1]
    if head_tok[0] != 'ident':
        raise QuilParseError('cannot parse {!r}'.format(text))
    if head in ('DEFGATE', 'DEFCIRCUIT', 'CONVERT'):
        raise ValueError('Unknown pyQuil instruction: {}'.format(text))

    if head in _NULLARY or head in ('RESET', 'LABEL', 'JUMP', 'JUMP-WHEN', 'JUMP-UNLESS',
                                    'MEASURE', 'DECLARE', 'PRAGMA', 'INCLUDE', 'TRUE', 'FALSE',
                                    'NOT', 'NEG', 'LOAD', 'STORE') \
            or head in _BINARY or head in _COMPARE:
        args = _tokens(text)[1:]
        bad = [t for k, t in args if k == 'bad']
        if bad:
            raise QuilParseError('cannot parse {!r} in {!r}'.format(bad[0], text))
        if head in _NULLARY:
            _nargs(head, args, 0)
            return _NULLARY[head]()
        if head == 'RESET':
            return Reset(*[_qubit(t) for t in args])
        if head == 'LABEL':
            _nargs(head, args, 1)
            return Label(_label(args[0]))
        if head == 'JUMP':
            _nargs(head, args, 1)
            return Jump(_label(args[0]))
        if head in ('JUMP-WHEN', 'JUMP-UNLESS'):
            _nargs(head, args, 2)
            cls = JumpWhen if head == 'JUMP-WHEN' else JumpUnless
            return cls(_label(args[0]), _addr(args[1]))
        if head == 'MEASURE':
            if len(args) not in (1, 2):
                raise QuilParseError('MEASURE takes a qubit and an optional address')
            if len(args) == 1:
                return Measure(_qubit(args[0]))
            return Measure(_qubit(args[0]), _addr(args[1]))
        if head == 'DECLARE':
            return _declare(text, args)
        if head == 'PRAGMA':
            return _pragma(args, text)
        if head == 'INCLUDE':
            _nargs(head, args, 1)
            if args[0][0] != 'string':
                raise QuilParseError('INCLUDE takes a quoted file name')
            return Include(args[0][1][1:-1])
        if head in ('TRUE', 'FALSE'):
            _nargs(head, args, 1)
            return Move(_addr(args[0]), 1 if head == 'TRUE' else 0)
        if head in ('NOT', 'NEG'):
            _nargs(head, args, 1)
            return (Not if head == 'NOT' else Neg)(_addr(args[0]))
        if head == 'LOAD':
            _nargs(head, args, 3)
            return Load(_addr(args[0]), args[1][1], _addr(args[2]))
        if head == 'STORE':
            _nargs(head, args, 3)
            return Store(args[0][1], _addr(args[1]), _operand(args[2]))
        if head in _BINARY:
            _nargs(head, args, 2)
            if head == 'EXCHANGE':
                return Exchange(_addr(args[0]), _addr(args[1]))
            return _BINARY[head](_addr(args[0]), _operand(args[1]))
        _nargs(head, args, 3)
        return _COMPARE[head](_addr(args[0]), _addr(args[1]), _operand(args[2]))

    # a gate: NAME[(params)] qubit+
    name, params, rest = _split_params(text)
    qubits = [_qubit(t) for t in _tokens(rest)]
    if not qubits:
        raise QuilParseError('cannot parse {!r}: a gate needs at least one qubit'.format(text))
    if name in STDGATES:
        return STDGATES[name](*params).relabel(qubits)
    return Call(name, params, qubits)


def _declare(text: str, args) -> Declare:
    """DECLARE name TYPE [n] [SHARING region [OFFSET n TYPE]*]."""
    if len(args) < 2 or args[0][0] != 'ident':
        raise QuilParseError('cannot parse {!r}'.format(text))
    name = args[0][1]
    rest = list(args[1:])
    kind, tok = rest.pop(0)
    size = 1
    if kind == 'addr':                         # TYPE[n], no space
        mtype, n = tok.split('[')
        size = int(n.rstrip(']').strip())
    elif kind == 'ident':
        mtype = tok
        if rest and rest[0][0] == 'addr' and rest[0][1].startswith('['):
            size = int(rest.pop(0)[1].strip('[] '))
    else:
        raise QuilParseError('cannot parse {!r}'.format(text))
    shared, offsets = None, None
    if rest:
        if rest[0] != ('ident', 'SHARING') or len(rest) < 2:
            raise QuilParseError('cannot parse {!r}'.format(text))
        shared = rest[1][1]
        rest = rest[2:]
        while rest:
            if rest[0] != ('ident', 'OFFSET') or len(rest) < 3 or rest[1][0] != 'number':
                raise QuilParseError('cannot parse {!r}'.format(text))
            offsets = (offsets or []) + [(int(rest[1][1]), rest[2][1])]
            rest = rest[3:]
    return Declare(name, mtype, size, shared, offsets)


def _pragma(args, text: str) -> Pragma:
    """PRAGMA name (identifier | integer)* ["freeform"]."""
    if not args or args[0][0] != 'ident':
        raise QuilParseError('cannot parse {!r}'.format(text))
    command = args[0][1]
    pargs: List[Any] = []
    freeform = None
    for k, (kind, tok) in enumerate(args[1:]):
        if kind == 'string':
            if k != len(args) - 2:
                raise QuilParseError('the freeform string ends a PRAGMA')
            freeform = tok[1:-1]
        elif kind == 'number' and re.fullmatch(r'\d+', tok):
            pargs.append(int(tok))
        elif kind == 'ident':
            pargs.append(tok)
        else:
            raise QuilParseError('cannot parse {!r}'.format(text))
    return Pragma(command, pargs or None, freeform)


def quil_to_program(quil: str) -> Program:
    """Parse Quil text into a Program (forest/__init__.py:216-219)."""
    return Program([_instruction(line) for line in _split_instructions(quil)])


# ---------------------------------------------------------------------------

class QuantumFlowQVM:
    """A quantum virtual machine over the GPU state, with the reference's
    status machine: 'connected' -> load -> 'loaded' -> run -> 'running' ->
    wait -> 'done' (forest/__init__.py:369-444).  ``load`` takes a Program,
    or anything whose ``str`` is Quil text (as the reference takes a pyQuil
    program); ``wavefunction`` returns the final State."""

    def __init__(self) -> None:
        self.program: Any = None
        self.status = 'connected'
        self._prog: Program = None
        self._ket: State = None

    def load(self, binary) -> 'QuantumFlowQVM':
        assert self.status in ['connected', 'done']
        prog = binary if isinstance(binary, Program) else quil_to_program(str(binary))
        self._prog = prog
        self.program = binary
        self.status = 'loaded'
        return self

    def write_memory(self, *, region_name: str, offset: int = 0,
                     value: int = None) -> 'QuantumFlowQVM':
        raise NotImplementedError()

    def run(self) -> 'QuantumFlowQVM':
        """Run the loaded program.  The status stays 'running' until
        ``wait``, as the reference keeps it for its callers."""
        assert self.status in ['loaded']
        self.status = 'running'
        self._ket = self._prog.run()
        return self

    def wait(self) -> 'QuantumFlowQVM':
        assert self.status == 'running'
        self.status = 'done'
        return self

    def read_from_memory_region(self, *, region_name: str,
                                offsets: Sequence[int] = None) -> Sequence[Any]:
        """Every value in the named region of the final state's memory, in
        the memory's order."""
        assert self.status == 'done'
        if offsets is not None:
            raise NotImplementedError('Offsets not yet supported')
        reg = Register(region_name)
        assert self._ket is not None
        return [value for addr, value in self._ket.memory.items() if addr.register == reg]

    def wavefunction(self) -> State:
        assert self.status == 'done'
        assert self._ket is not None
        return self._ket

"""Hybrid quantum-classical programs (quantumflow/programs.py:1-560).

A ``Program`` is a list of instructions run under a program counter kept
in the state's classical memory, following the Quil instruction set:
labels and (conditional) jumps, halting, classical memory declarations,
named-gate calls, and any Operation -- gates, measurements, resets,
classical arithmetic -- as an instruction.  The quantum state stays a
device tensor throughout: gates run on the GPU kernels, ``Measure`` draws
its outcome from one reduction kernel's P(|0>) and collapses the state
with one diagonal-gate pass, and the classical control reads only that
outcome on the host.  ``Program.evolve`` runs the same loop on a Density.
"""

from typing import Dict, List, Sequence, Tuple, Union

from .cbits import Addr, Register
from .qubits import Qubits
from .states import State, zero_state
from .stdgates import STDGATES

__all__ = ['Instruction', 'Program', 'DefCircuit', 'Wait', 'Nop', 'Halt', 'Label', 'Jump',
           'JumpWhen', 'JumpUnless', 'Pragma', 'Include', 'Call', 'Declare', 'Load',
           'Store', 'PC', 'TARGETS', 'NAMEDGATES', 'HALTED']

# Private register holding the program's own state (programs.py:44-49)
_PROG_STATE = Register('_prog_state_')
PC = _PROG_STATE['pc']
NAMEDGATES = _PROG_STATE['namedgates']
TARGETS = _PROG_STATE['targets']
WAIT = _PROG_STATE['wait']

HALTED = -1             # program counter of a finished program


class Instruction:
    """One instruction of a program (programs.py:54-90)."""

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

    def quil(self) -> str:
        return self.name

    def __str__(self) -> str:
        return self.quil()

    def run(self, ket: State) -> State:
        raise NotImplementedError()

    def evolve(self, rho):
        """A purely classical instruction acts on a Density as on a state."""
        return self.run(rho)


class Program(Instruction):
    """A program for a hybrid quantum computer (programs.py:93-171)."""

    def __init__(self, instructions: List = None, name: str = None,
                 params: dict = None) -> None:
        self.instructions = list(instructions) if instructions is not None else []

    def quil(self) -> str:
        return '\n'.join([str(i) for i in self.instructions] + [''])

    @property
    def qubits(self) -> Qubits:
        qubits = {q for instr in self.instructions for q in instr.qubits}
        return sorted(qubits)

    def __iadd__(self, other) -> 'Program':
        self.instructions.append(other)
        return self

    def __len__(self) -> int:
        return len(self.instructions)

    def __getitem__(self, key: int):
        return self.instructions[key]

    def __iter__(self):
        return iter(self.instructions)

    def _initialize(self, state):
        """PC 0, the label table and the named gates (programs.py:136-148)."""
        targets = {instr.target: pc for pc, instr in enumerate(self.instructions)
                   if isinstance(instr, Label)}
        return state.update({PC: 0, TARGETS: targets, NAMEDGATES: STDGATES.copy()})

    def _loop(self, state, step):
        state = self._initialize(state)
        pc = 0
        n = len(self.instructions)
        while 0 <= pc < n:
            state = state.update({PC: pc + 1})
            state = step(self.instructions[pc], state)
            pc = state.memory[PC]
        return state

    def run(self, ket: State = None) -> State:
        """Run from ``ket`` (default |0...0> on the program's qubits):
        each instruction sees PC already advanced, so a jump overwrites it
        (programs.py:152-170)."""
        if ket is None:
            ket = zero_state(self.qubits)
        return self._loop(ket, lambda instr, k: instr.run(k))

    def evolve(self, rho=None):
        """The same loop on a density (programs.py:172-188)."""
        if rho is None:
            rho = zero_state(self.qubits).asdensity()
        return self._loop(rho, lambda instr, r: instr.evolve(r))


class DefCircuit(Program):
    """A named, parameterised sub-program; recorded and printed only, as in
    the reference (programs.py:193-245)."""

    def __init__(self, name: str, params: Dict[str, float], qubits: Qubits = None,
                 instructions: List = None) -> None:
        super().__init__(instructions)
        self.progname = name
        self.params = params
        self._qubits = list(qubits) if qubits else []

    @property
    def qubits(self) -> Qubits:
        return self._qubits

    def quil(self) -> str:
        fparams = '(' + ','.join(map(str, self.params)) + ')' if self.params else ''
        fqubits = ' ' + ' '.join(map(str, self.qubits)) if self.qubits else ''
        out = '{} {}{}{}:\n'.format(self.name, self.progname, fparams, fqubits)
        for instr in self.instructions:
            out += '    ' + str(instr) + '\n'
        return out


class Wait(Instruction):
    """Hand control to a higher level; a no-op here (programs.py:248-252)."""

    def run(self, ket: State) -> State:
        return ket


class Nop(Instruction):
    """No operation (programs.py:255-258)."""

    def run(self, ket: State) -> State:
        return ket


class Halt(Instruction):
    """Stop the program (programs.py:261-265)."""

    def run(self, ket: State) -> State:
        return ket.update({PC: HALTED})


class Load(Instruction):
    """Quil LOAD; printed but not executable, as in the reference
    (programs.py:268-281)."""

    def __init__(self, target: Addr, left: str, right: Addr) -> None:
        self.target, self.left, self.right = target, left, right

    def quil(self) -> str:
        return '{} {} {} {}'.format(self.name, self.target, self.left, self.right)

    def run(self, ket: State) -> State:
        raise NotImplementedError()


class Store(Instruction):
    """Quil STORE; printed but not executable, as in the reference
    (programs.py:284-298)."""

    def __init__(self, target: str, left: Addr, right: Union[int, Addr]) -> None:
        self.target, self.left, self.right = target, left, right

    def quil(self) -> str:
        return '{} {} {} {}'.format(self.name, self.target, self.left, self.right)

    def run(self, ket: State) -> State:
        raise NotImplementedError()


class Label(Instruction):
    """A jump target (programs.py:301-310)."""

    def __init__(self, target: str) -> None:
        self.target = target

    def quil(self) -> str:
        return '{} @{}'.format(self.name, self.target)

    def run(self, ket: State) -> State:
        return ket


class Jump(Instruction):
    """Unconditional jump to a label (programs.py:313-323)."""

    def __init__(self, target: str) -> None:
        self.target = target

    def quil(self) -> str:
        return '{} @{}'.format(self.name, self.target)

    def run(self, ket: State) -> State:
        return ket.update({PC: ket.memory[TARGETS][self.target]})


class JumpWhen(Instruction):
    """Jump to a label when a classical cell is non-zero (programs.py:326-343)."""

    def __init__(self, target: str, condition: Addr) -> None:
        self.target = target
        self.condition = condition

    @property
    def name(self) -> str:
        return 'JUMP-WHEN'

    def quil(self) -> str:
        return '{} @{} {}'.format(self.name, self.target, self.condition)

    def run(self, ket: State) -> State:
        if ket.memory[self.condition]:
            return ket.update({PC: ket.memory[TARGETS][self.target]})
        return ket


class JumpUnless(Instruction):
    """Jump to a label when a classical cell is zero (programs.py:346-364)."""

    def __init__(self, target: str, condition: Addr) -> None:
        self.target = target
        self.condition = condition

    @property
    def name(self) -> str:
        return 'JUMP-UNLESS'

    def quil(self) -> str:
        return '{} @{} {}'.format(self.name, self.target, self.condition)

    def run(self, ket: State) -> State:
        if not ket.memory[self.condition]:
            return ket.update({PC: ket.memory[TARGETS][self.target]})
        return ket


class Pragma(Instruction):
    """PRAGMA <command> <arg>* "<freeform>"?; no effect on the state
    (programs.py:367-391)."""

    def __init__(self, command: str, args: Sequence = None, freeform: str = None) -> None:
        self.command = command
        self.args = args
        self.freeform = freeform

    def quil(self) -> str:
        out = ['PRAGMA {}'.format(self.command)]
        if self.args:
            out.extend(str(a) for a in self.args)
        if self.freeform:
            out.append('"{}"'.format(self.freeform))
        return ' '.join(out)

    def run(self, ket: State) -> State:
        return ket


class Include(Instruction):
    """INCLUDE "<file>"; recorded, not executable (programs.py:394-408)."""

    def __init__(self, filename: str, program: Program = None) -> None:
        self.filename = filename
        self.program = program

    def quil(self) -> str:
        return '{} "{}"'.format(self.name, self.filename)

    def run(self, ket: State) -> State:
        raise NotImplementedError()


class Call(Instruction):
    """Run a named gate from the program's gate table on the given qubits
    (programs.py:411-442); RuntimeError for an unknown name."""

    def __init__(self, name: str, params: Sequence, qubits: Qubits) -> None:
        self.gatename = name
        self.params = list(params)
        self._qubits = list(qubits)

    def quil(self) -> str:
        fqubits = ' ' + ' '.join(str(q) for q in self.qubits) if self.qubits else ''
        fparams = '(' + ', '.join(str(p) for p in self.params) + ')' if self.params else ''
        return '{}{}{}'.format(self.gatename, fparams, fqubits)

    def run(self, ket: State) -> State:
        namedgates = ket.memory[NAMEDGATES]
        if self.gatename not in namedgates:
            raise RuntimeError('Unknown named gate')
        gate = namedgates[self.gatename](*self.params).relabel(self.qubits)
        return gate.run(ket)

    def evolve(self, rho):
        namedgates = rho.memory[NAMEDGATES]
        if self.gatename not in namedgates:
            raise RuntimeError('Unknown named gate')
        gate = namedgates[self.gatename](*self.params).relabel(self.qubits)
        return gate.evolve(rho)


class Declare(Instruction):
    """DECLARE name type [size]: zero `size` cells of a new register
    (programs.py:489-522)."""

    def __init__(self, memory_name: str, memory_type: str, memory_size: int,
                 shared_region: str = None, offsets: List[Tuple[int, str]] = None) -> None:
        self.memory_name = memory_name
        self.memory_type = memory_type
        self.memory_size = memory_size
        self.shared_region = shared_region
        self.offsets = offsets

    def quil(self) -> str:
        out = ['DECLARE', self.memory_name, self.memory_type]
        if self.memory_size != 1:
            out.append('[{}]'.format(self.memory_size))
        if self.shared_region is not None:
            out += ['SHARING', self.shared_region]
        return ' '.join(out)

    def run(self, ket: State) -> State:
        reg = Register(self.memory_name, self.memory_type)
        return ket.update({reg[i]: 0 for i in range(self.memory_size)})

"""Quil text -> Program (quantumflow_amd.forest.quil_to_program) on the host:
the reference's parser known answers (tests/test_parser.py, tests/test_quil.py)
and its Quil fixture files (copied under tests/golden/quil/), plus
print -> parse round trips.  Nothing here runs a state."""
import cmath
import math
import os

import numpy as np
import pytest

import quantumflow_amd as qf
from quantumflow_amd.forest import QuilParseError, quil_to_program

HERE = os.path.dirname(os.path.abspath(__file__))
QUILDIR = os.path.join(HERE, 'golden', 'quil')

# tests/test_quil.py:12-22
QUIL_FILES = ['hello_world.quil', 'empty.quil', 'classical_logic.quil', 'control_flow.quil',
              'measure.quil', 'qaoa.quil', 'bell.quil']


def _read(name):
    with open(os.path.join(QUILDIR, name)) as f:
        return f.read()


def _same(quil, *instructions):
    """tests/test_parser.py:21-25: parsed and built programs print alike."""
    assert quil_to_program(quil).quil() == qf.Program(list(instructions)).quil()


def _arg(quil):
    return list(quil_to_program(quil)[0].params.values())[0]


def test_empty_and_simple():
    _same('')
    for quil in ('HALT', 'WAIT', 'NOP', 'RESET'):
        prog = quil_to_program(quil)
        assert len(prog) == 1 and prog[0].quil() == quil
    _same('RESET 1', qf.Reset(1))


def test_parameter_expressions():
    assert _arg('RX(1) 0') == 1
    assert _arg('RX(2.9) 0') == 2.9
    assert _arg('RX(-2.9) 0') == -2.9
    assert _arg('RX(+2.9) 0') == +2.9
    assert _arg('RX(+2.9/4.2) 0') == +2.9 / 4.2
    assert _arg('RX(+2.9*4.2) 0') == +2.9 * 4.2
    assert _arg('RX(2.9+4.2) 0') == 2.9 + 4.2
    assert _arg('RX(2.9-4.2) 0') == 2.9 - 4.2
    assert _arg('RX(pi) 0') == math.pi
    assert _arg('RX(2.0*pi) 0') == 2.0 * math.pi
    assert _arg('RX(SIN(1.0)) 0') == math.sin(1.0)
    assert _arg('RX(EXP(3.3)) 0') == math.exp(3.3)
    assert abs(math.cos(2.3) - _arg('RX(COS(2.3)) 0')) < 1e-12
    assert abs(math.sqrt(42) - _arg('RX( SQRT( 42  )   ) 0')) < 1e-12
    assert cmath.isclose(cmath.exp(2j), _arg('RX(CIS(2)) 0'))
    assert _arg('RX(2.3 i) 0') == 2.3j
    assert abs(math.exp(2) - _arg('RX(EXP(2)) 0')) < 1e-12
    assert abs(4 - _arg('RX(2^2) 0')) < 1e-12
    # precedence: a sign binds tighter than ^, ^ is right-associative and
    # tighter than * and /, which are tighter than + and -
    assert _arg('RX(-2^2) 0') == 4
    assert _arg('RX(2^3^2) 0') == 2 ** 9
    assert _arg('RX(1+2*3^2) 0') == 19
    assert _arg('RX((1+2)*3) 0') == 9
    assert _arg('RX(-pi/2) 0') == -math.pi / 2
    assert _arg('RX(1/2/4) 0') == 0.125
    with pytest.raises(QuilParseError):
        quil_to_program('RX(%theta) 0')
    with pytest.raises(QuilParseError):
        quil_to_program('RX(2*) 0')
    with pytest.raises(QuilParseError):
        quil_to_program('RX((1) 0')


def test_classical():
    ro = qf.Register()
    cmd = quil_to_program('TRUE ro[1]')[0]
    assert isinstance(cmd, qf.Move) and cmd.target == ro[1] and cmd.quil() == 'MOVE ro[1] 1'
    b = qf.Register('b')
    cmd = quil_to_program('FALSE b[2]')[0]
    assert isinstance(cmd, qf.Move) and cmd.target == b[2] and cmd.quil() == 'MOVE b[2] 0'
    for quil, cls in (('NOT b[3]', qf.Not), ('NEG b[3]', qf.Neg)):
        cmd = quil_to_program(quil)[0]
        assert isinstance(cmd, cls) and cmd.target == b[3] and cmd.quil() == quil
    for quil, cls in (('AND b[0] b[1]', qf.And), ('IOR b[0] b[1]', qf.Ior),
                      ('XOR b[0] b[1]', qf.Xor), ('MOVE b[0] b[1]', qf.Move),
                      ('EXCHANGE b[0] b[1]', qf.Exchange), ('ADD b[0] b[1]', qf.Add),
                      ('MUL b[0] b[1]', qf.Mul), ('SUB b[0] b[1]', qf.Sub),
                      ('DIV b[0] b[1]', qf.Div), ('ADD b[0] 4', qf.Add), ('MUL b[0] 2', qf.Mul),
                      ('SUB b[0] 3', qf.Sub), ('DIV b[0] 2', qf.Div),
                      ('LOAD b[0] this b[1]', qf.Load), ('STORE that b[0] b[1]', qf.Store),
                      ('EQ a[1] c[0] b[1]', qf.EQ), ('GT a[1] c[0] b[1]', qf.GT),
                      ('GE a[1] c[0] b[1]', qf.GE), ('LT a[1] c[0] b[1]', qf.LT),
                      ('LE a[1] c[0] b[1]', qf.LE)):
        cmd = quil_to_program(quil)[0]
        assert isinstance(cmd, cls) and cmd.quil() == quil, quil
    a = qf.Register('a')
    _same('TRUE b[0]', qf.Move(b[0], 1))
    _same('FALSE b[0]', qf.Move(b[0], 0))
    _same('NOT b[0]', qf.Not(b[0]))
    _same('AND b[0] a[1]', qf.And(b[0], a[1]))
    _same('XOR b[0] b[1]', qf.Xor(b[0], b[1]))
    _same('MOVE b[0] b[1]', qf.Move(b[0], b[1]))
    _same('EXCHANGE b[0] b[1]', qf.Exchange(b[0], b[1]))
    # memory references: a bare name is offset 0, [n] is ro[n]; every cell
    # gets the register's default BIT type, as the reference's conversion
    assert quil_to_program('NOT flag')[0].target == qf.Register('flag')[0]
    assert quil_to_program('NOT [4]')[0].target == ro[4]
    assert quil_to_program('MOVE b[0] 1.5')[0].source == 1.5
    assert quil_to_program('ADD b[0] -4')[0].source == -4
    with pytest.raises(ValueError):
        quil_to_program('CONVERT b[0] b[1]')


def test_declare():
    for quil in ('DECLARE ro BITS', 'DECLARE ro BITS [10]', 'DECLARE ro BITS [10] SHARING rs'):
        cmd = quil_to_program(quil)[0]
        assert isinstance(cmd, qf.Declare) and cmd.quil() == quil
    cmd = quil_to_program('DECLARE theta REAL[4]')[0]
    assert (cmd.memory_name, cmd.memory_type, cmd.memory_size) == ('theta', 'REAL', 4)
    cmd = quil_to_program('DECLARE ro BIT [10] SHARING rs OFFSET 2 REAL')[0]
    assert cmd.shared_region == 'rs' and cmd.offsets == [(2, 'REAL')]


def test_labels_jumps_measure_pragma():
    cmd = quil_to_program('LABEL @some_target')[0]
    assert isinstance(cmd, qf.Label) and cmd.target == 'some_target'
    assert cmd.quil() == 'LABEL @some_target'
    cmd = quil_to_program('JUMP @some_target')[0]
    assert isinstance(cmd, qf.Jump) and cmd.target == 'some_target'
    for quil, cls in (('JUMP-UNLESS @some_target ro[2]', qf.JumpUnless),
                      ('JUMP-WHEN @some_target b[3]', qf.JumpWhen)):
        cmd = quil_to_program(quil)[0]
        assert isinstance(cmd, cls) and cmd.quil() == quil
    cb = qf.Register('cb')
    _same('LABEL @test_1', qf.Label('test_1'))
    _same('JUMP @test_1', qf.Jump('test_1'))
    _same('JUMP-WHEN @test_1 cb[0]', qf.JumpWhen('test_1', cb[0]))
    _same('JUMP-UNLESS @test_1 cb[1]', qf.JumpUnless('test_1', cb[1]))
    for quil in ('MEASURE 1', 'MEASURE 3 reg0[2]'):
        cmd = quil_to_program(quil)[0]
        assert isinstance(cmd, qf.Measure) and cmd.quil() == quil
    _same('MEASURE 0', qf.Measure(0))
    _same('MEASURE 0 b[1]', qf.Measure(0, qf.Register('b')[1]))
    for quil, nargs, free in (('PRAGMA somename', 0, None), ('PRAGMA somename arg0', 1, None),
                              ('PRAGMA somename arg0 arg1', 2, None),
                              ('PRAGMA somename "some string"', 0, 'some string'),
                              ('PRAGMA somename arg2 arg3 arg7 "some string"', 3, 'some string')):
        cmd = quil_to_program(quil)[0]
        assert isinstance(cmd, qf.Pragma) and cmd.quil() == quil
        assert len(cmd.args or []) == nargs and cmd.freeform == free
    _same('PRAGMA gate_time H "10 ns"', qf.Pragma('gate_time', ['H'], '10 ns'))
    _same('PRAGMA qubit 0', qf.Pragma('qubit', [0]))
    _same('PRAGMA NO-NOISE', qf.Pragma('NO-NOISE'))


def test_gates():
    for g in ('I', 'H', 'X', 'Y', 'Z', 'S', 'T'):
        cmd = quil_to_program('{} 42'.format(g))[0]
        assert cmd.name == g and cmd.qubits == (42,) and not cmd.params
    for g in ('SWAP', 'CZ', 'CNOT', 'ISWAP'):
        cmd = quil_to_program('{} 2 42'.format(g))[0]
        assert cmd.name == g and cmd.qubits == (2, 42) and not cmd.params
    for g in ('CCNOT', 'CSWAP'):
        cmd = quil_to_program('{} 2 42 5'.format(g))[0]
        assert cmd.name == g and cmd.qubits == (2, 42, 5) and not cmd.params
    for g in ('RX', 'RY', 'RZ'):
        cmd = quil_to_program('{}(2.8) 42'.format(g))[0]
        assert cmd.name == g and cmd.qubits == (42,) and cmd.params == {'theta': 2.8}
    for g in ('CPHASE00', 'CPHASE01', 'CPHASE10', 'CPHASE', 'PSWAP'):
        cmd = quil_to_program('{}(0.5) 2 42'.format(g))[0]
        assert cmd.name == g and cmd.qubits == (2, 42) and cmd.params == {'theta': 0.5}
    _same('RX(123) 0', qf.Call('RX', [123], qubits=[0]))
    _same('CPHASE00(0) 0 1', qf.Call('CPHASE00', [0], [0, 1]))
    _same('A(8,9) 0', qf.Call('A', [8, 9], [0]))
    _same('A(8, 9) 0', qf.Call('A', [8, 9], [0]))
    # a standard gate parsed from Quil has the same matrix as one built directly
    cmd = quil_to_program('RX(pi/2) 3')[0]
    np.testing.assert_array_equal(cmd.matrix, qf.RX(math.pi / 2, 3).matrix)
    with pytest.raises(QuilParseError):
        quil_to_program('H')
    with pytest.raises(QuilParseError):
        quil_to_program('H q')
    with pytest.raises(ValueError):
        quil_to_program('DEFGATE FOO:')


def test_comments_and_separators():
    prog = quil_to_program('# a comment\nH 0 # trailing\n\nX 1; CNOT 0 1\nPRAGMA x "a # b; c"')
    assert [i.quil() for i in prog] == ['H 0', 'X 1', 'CNOT 0 1', 'PRAGMA x "a # b; c"']


def test_quil_files():
    """tests/test_quil.py:25-48: every fixture parses; BELL does not."""
    sizes = {}
    for name in QUIL_FILES:
        sizes[name] = len(quil_to_program(_read(name)))
    assert sizes == {'hello_world.quil': 2, 'empty.quil': 0, 'classical_logic.quil': 6,
                     'control_flow.quil': 7, 'measure.quil': 4, 'qaoa.quil': 11,
                     'bell.quil': 2}
    inc = quil_to_program(_read('include.quil'))
    assert len(inc) == 1 and isinstance(inc[0], qf.Include) and inc[0].filename == 'bell.quil'
    with pytest.raises(RuntimeError):
        quil_to_program(_read('unparsable.quil'))


def test_print_parse_round_trip():
    """A program's Quil text parses back into a program that prints the
    same, for gates with exact (repr-printed) parameters, classical ops and
    control flow."""
    rng = np.random.default_rng(3)
    ro = qf.Register()
    k = qf.Register('k', 'INT')
    prog = qf.Program([qf.Declare('k', 'INT', 2), qf.Label('top')])
    for _ in range(40):
        r = rng.uniform()
        if r < 0.4:
            prog += qf.Call('RX', [float(rng.uniform(-7, 7))], [int(rng.integers(6))])
        elif r < 0.7:
            a, b = (int(x) for x in rng.choice(6, 2, replace=False))
            prog += qf.Call('CNOT', [], [a, b])
        elif r < 0.8:
            prog += qf.Measure(int(rng.integers(6)), ro[int(rng.integers(4))])
        else:
            prog += qf.Add(k[0], int(rng.integers(1, 5)))
    prog += qf.JumpWhen('top', ro[0])
    prog += qf.Halt()
    text = prog.quil()
    again = quil_to_program(text)
    assert len(again) == len(prog)
    for a, b in zip(prog, again):
        if isinstance(a, qf.Call):
            # a standard-gate call comes back as that gate, its parameters
            # bit for bit (the gate itself prints them to 10 digits)
            assert b.name == a.gatename and list(b.qubits) == a.qubits
            assert list(b.params.values()) == a.params
        else:
            assert b.quil() == a.quil()


def test_qvm_status_machine():
    qvm = qf.forest.QuantumFlowQVM()
    assert qvm.status == 'connected'
    with pytest.raises(AssertionError):
        qvm.run()
    qvm.load('H 0\nCNOT 0 1')
    assert qvm.status == 'loaded' and len(qvm._prog) == 2
    with pytest.raises(AssertionError):
        qvm.load('X 0')
    with pytest.raises(AssertionError):
        qvm.wait()
    with pytest.raises(NotImplementedError):
        qvm.write_memory(region_name='ro')
    with pytest.raises(AssertionError):
        qvm.read_from_memory_region(region_name='ro')

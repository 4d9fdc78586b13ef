"""CPU checks of the C-ABI boundary: libqfx.so loads and exports every
function include/qfx.h declares, the ctypes table covers exactly those, and
error codes map onto the reference's exception types.  No kernel runs."""
import ctypes
import os
import subprocess

import pytest

from quantumflow_amd import _native


def test_library_built():
    assert os.path.exists(_native.LIB_PATH), "run python -m quantumflow_amd._build"


def test_exports_every_declared_symbol():
    declared = _native.header_symbols()
    assert len(declared) >= 15
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    # the ctypes binding declares exactly the header's functions
    assert sorted(_native.SIGNATURES) == declared


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(['nm', '-D', '--defined-only', _native.LIB_PATH],
                         capture_output=True, text=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if ' T ' in line}
    for name in _native.header_symbols():
        assert name in syms, name        # unmangled


def test_version_and_errors_without_gpu():
    lib = _native.load()
    assert b'gfx950' in lib.qfx_version()
    # argument validation happens on the host before any HIP call
    rc = lib.qfx_apply_matrix(None, None, None, 4, 1, None, None)
    assert rc == _native.QFX_EINVAL or rc == _native.QFX_ERANGE
    rc = lib.qfx_apply_matrix(None, ctypes.c_void_p(16), ctypes.c_void_p(16), 40, 1,
                              _native.int_array([0]), _native.c128_array([1, 0, 0, 1]))
    assert rc == _native.QFX_ERANGE
    rc = lib.qfx_apply_diagonal(None, ctypes.c_void_p(16), ctypes.c_void_p(16), 4, 2,
                                _native.int_array([1, 1]), _native.c128_array([1] * 4))
    assert rc == _native.QFX_EINVAL
    assert 'repeated' in _native.last_error()
    rc = lib.qfx_apply_monomial(None, ctypes.c_void_p(16), ctypes.c_void_p(16), 4, 1,
                                _native.int_array([0]), _native.int_array([0, 0]), None)
    assert rc == _native.QFX_EINVAL
    with pytest.raises(ValueError):
        _native.check(_native.QFX_EINVAL, 'x')
    with pytest.raises(RuntimeError):
        _native.check(_native.QFX_EHIP, 'x')


def test_workspace_size_formula():
    lib = _native.load()
    assert lib.qfx_reduce_workspace_bytes(1 << 28) >= (1 << 12) * 17 * 8
    assert lib.qfx_reduce_workspace_bytes(4) >= 1024 * 16

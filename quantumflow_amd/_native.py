"""ctypes binding of libqfx.so, the C-ABI declared in include/qfx.h.

The library is required: there is no CPU fallback.  Importing this module
succeeds without the library (so host-only code and the CPU test suite can
run), but the first call raises ``RuntimeError`` naming the missing file.
"""

import ctypes
import os
import re
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libqfx.so')
HEADER_DIR = os.path.join(os.path.dirname(HERE), 'include')

QFX_OK = 0
QFX_EINVAL = -1
QFX_ERANGE = -2
QFX_EHIP = -3
QFX_ENODEV = -4
QFX_EKERNEL = -5


class QfxC128(ctypes.Structure):
    _fields_ = [('re', ctypes.c_double), ('im', ctypes.c_double)]


class QfxTileOp(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('k', ctypes.c_int32),
                ('coef', ctypes.c_int32), ('src', ctypes.c_int32),
                ('t', ctypes.c_int32 * 5), ('sorted', ctypes.c_int32 * 5),
                ('active', ctypes.c_uint32), ('phase', ctypes.c_uint32),
                ('lds', ctypes.c_uint32 * 10)]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_ip = ctypes.POINTER(ctypes.c_int)
_cp = ctypes.POINTER(QfxC128)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every symbol declared in include/qfx.h
SIGNATURES = {
    'qfx_last_error': (ctypes.c_char_p, []),
    'qfx_version': (ctypes.c_char_p, []),
    'qfx_init': (_i, [_i64p]),
    'qfx_apply_matrix': (_i, [_vp, _vp, _vp, _i, _i, _ip, _cp]),
    'qfx_apply_matrix_dev': (_i, [_vp, _vp, _vp, _i, _i, _ip, _vp]),
    'qfx_apply_diagonal': (_i, [_vp, _vp, _vp, _i, _i, _ip, _cp]),
    'qfx_apply_monomial': (_i, [_vp, _vp, _vp, _i, _i, _ip, _ip, _cp]),
    'qfx_apply_diagonal_inplace': (_i, [_vp, _vp, _i, _i, _ip, _cp]),
    'qfx_apply_monomial_inplace': (_i, [_vp, _vp, _i, _i, _ip, _ip, _cp]),
    'qfx_reduce_workspace_bytes': (_i64, [_i64]),
    'qfx_inner': (_i, [_vp, _vp, _vp, _vp, _i64, _vp]),
    'qfx_abs': (_i, [_vp, _vp, _vp, _i64]),
    'qfx_abs2': (_i, [_vp, _vp, _vp, _i64]),
    'qfx_expect_diag': (_i, [_vp, _vp, _vp, _vp, _i64, _vp]),
    'qfx_expect_diag_noalloc': (_i, [_vp, _vp, _vp, _vp, _i64]),
    'qfx_expect_z': (_i, [_vp, _vp, _vp, _i, _vp]),
    'qfx_norm2': (_i, [_vp, _vp, _vp, _i64, _vp]),
    'qfx_qubit_probs': (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    'qfx_permute_qubits': (_i, [_vp, _vp, _vp, _i, _ip]),
    'qfx_zero_state': (_i, [_vp, _vp, _i]),
    'qfx_plus_state': (_i, [_vp, _vp, _i]),
    'qfx_sample_workspace_bytes': (_i64, [_i]),
    'qfx_sample': (_i, [_vp, _vp, _vp, _i, _i64, ctypes.c_uint64, _vp]),
    'qfx_histogram': (_i, [_vp, _vp, _vp, _i64, _i]),
    'qfx_density_outer': (_i, [_vp, _vp, _vp, _i]),
    'qfx_density_identity': (_i, [_vp, _vp, _i, ctypes.c_double]),
    'qfx_density_diag': (_i, [_vp, _vp, _vp, _i]),
    'qfx_density_trace': (_i, [_vp, _vp, _vp, _i]),
    'qfx_density_axpy': (_i, [_vp, _vp, _vp, ctypes.c_double, _i64]),
    'qfx_density_partial_trace': (_i, [_vp, _vp, _vp, _i, ctypes.c_uint32]),
    'qfx_cut_table': (_i, [_vp, _vp, _i, _vp, _vp, _i]),
    'qfx_cut_counts': (_i, [_vp, _vp, _i, _vp, _i]),
    'qfx_qaoa_check': (_i, [_vp, _i, _vp, _i, _vp, _vp, _i, _i, _i64, _i64, _i64, _i64, _i64, _i,
                            _i]),
    'qfx_qaoa_run': (_i, [_vp, _vp]),
    'qfx_qaoa_status': (_i, [_vp, _vp]),
    'qfx_mlp_gemm': (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i64, _i64, _i64, _i64, _vp, _i]),
    'qfx_mlp_relu_bwd': (_i, [_vp, _vp, _vp, _vp, _i64]),
    'qfx_mlp_colsum': (_i, [_vp, _vp, _vp, _i, _i]),
    'qfx_mlp_mse_workspace': (_i64, [_i64]),
    'qfx_mlp_mse': (_i, [_vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    'qfx_mlp_adam': (_i, [_vp, _vp, _vp, _vp, _vp, _i64, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, ctypes.c_float, _i, ctypes.c_float]),
    'qfx_tile_prepare_ops': (_i, [_vp, _i]),
    'qfx_tile_check_ops': (_i, [_i, _i, _i, _vp, _i, _i, _ip, _i]),
    'qfx_tile_pass': (_i, [_vp, _vp, _i, _i, _ip, _ip, _vp, _i, _vp, _i, _vp, _i, _i]),
}

_lib = None
_lock = threading.Lock()


def header_symbols():
    """Names of every function declared in include/*.h."""
    names = []
    for fn in sorted(os.listdir(HEADER_DIR)):
        if not fn.endswith('.h'):
            continue
        text = open(os.path.join(HEADER_DIR, fn)).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        names += re.findall(r'\b(qfx_\w+)\s*\(', text)
    return sorted(set(names))


def load():
    """Load libqfx.so and bind every signature; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                'libqfx.so not found at %s: build it with '
                '"python -m quantumflow_amd._build" (the HIP path has no CPU '
                'fallback)' % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def last_error():
    return load().qfx_last_error().decode('utf-8', 'replace')


def check(rc, what=''):
    """Map a qfx return code onto the reference's exception types."""
    if rc == QFX_OK:
        return
    msg = '%s: %s' % (what, last_error()) if what else last_error()
    if rc in (QFX_EINVAL, QFX_ERANGE):
        raise ValueError(msg)
    raise RuntimeError('qfx error %d: %s' % (rc, msg))


def call(name, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc


def int_array(values):
    values = list(values)
    return (ctypes.c_int * max(1, len(values)))(*values)


def c128_array(values):
    values = list(values)
    arr = (QfxC128 * max(1, len(values)))()
    for i, z in enumerate(values):
        arr[i].re = float(z.real)
        arr[i].im = float(z.imag)
    return arr

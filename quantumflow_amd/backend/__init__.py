"""QuantumFlow tensor-library backend (quantumflow/backend/__init__.py:1-39).

One backend module is bound at import from QUANTUMFLOW_BACKEND; this package
ships the gfx950 HIP backend ``rocm`` (quantumflow_amd.backend.rocmbk).
"""
from ..config import BACKEND, SEED  # noqa: F401

if BACKEND == 'rocm':
    from .rocmbk import *  # noqa: F401,F403
    from .rocmbk import (  # noqa: F401  (names star-import skips)
        __version__, name, version, pi, GateKind, classify, apply_gate,
        as_device, stream_handle, workspace, probabilities, expectation,
        expectation_z, norm, zero_state_tensor, qubit_probs, draw_seed,
        sample_indices, sample_counts, sample_expectation)
else:  # pragma: no cover  (config.py already rejected the name)
    raise ValueError('Unknown backend: ' + BACKEND)

__all__ = [
    'BKTensor', 'CTYPE', 'DEVICE', 'FTYPE', 'MAX_QUBITS', 'TENSOR',
    'TL', 'TensorLike', 'absolute', 'arccos', 'astensor',
    'ccast', 'cis', 'conj', 'cos', 'diag', 'evaluate', 'exp', 'fcast',
    'gpu_available', 'imag', 'inner', 'minimum',
    'outer', 'matmul',
    'rank', 'real', 'reshape', 'set_random_seed', 'sin',
    'sqrt', 'sum', 'tensormul', 'trace', 'transpose',
    'getitem', 'astensorproduct', 'productdiag',
    'EINSUM_SUBSCRIPTS', 'einsum',
    '__version__', '__name__']

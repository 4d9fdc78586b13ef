"""QuantumFlow configuration for the MI355X engine.

Mirrors quantumflow/config.py:19-62: environment variables with the
``QUANTUMFLOW_`` prefix, a closed tuple of backend names validated at import
(ValueError on an unknown name, config.py:47-49), the comparison TOLERANCE
and an optional random seed.

Backend names: ``rocm`` is this package's gfx950 HIP backend (the default
and only backend it ships; the reference's numpy/torch/tensorflow backends
are not part of this package).

Extra knobs (same prefix):
  QUANTUMFLOW_DEVICE        torch device string, default 'cuda'
  QUANTUMFLOW_FUSION        0 disables the tiled/fused circuit executor
  QUANTUMFLOW_TILE_BITS     override the executor's tile width (qubits)
  QUANTUMFLOW_AUTOTUNE      path of the autotune cache (JSON)
"""

import logging
import os
import random

_PREFIX = 'QUANTUMFLOW_'

version = '0.1.0'

logging.getLogger('quantumflow').addHandler(logging.StreamHandler())
_LOGLEVEL = os.getenv(_PREFIX + 'LOG', None)
if _LOGLEVEL is not None:
    logging.getLogger('quantumflow').setLevel(_LOGLEVEL)

DEFAULT_BACKEND = 'rocm'
BACKENDS = ('rocm',)

BACKEND = os.getenv(_PREFIX + 'BACKEND', DEFAULT_BACKEND)
if BACKEND not in BACKENDS:
    raise ValueError('Unknown backend: {}BACKEND={}'.format(_PREFIX, BACKEND))
logging.getLogger('quantumflow').info("QuantumFlow Backend: %s", BACKEND)

TOLERANCE = 1e-6
"""Tolerance used in various floating point comparisons"""

_ENVSEED = os.getenv(_PREFIX + 'SEED', None)
SEED = int(_ENVSEED) if _ENVSEED is not None else None
if SEED is not None:
    random.seed(SEED)

DEVICE = os.getenv(_PREFIX + 'DEVICE', 'cuda')
FUSION = os.getenv(_PREFIX + 'FUSION', '1') != '0'
TILE_BITS = int(os.getenv(_PREFIX + 'TILE_BITS', '0'))
AUTOTUNE_CACHE = os.getenv(
    _PREFIX + 'AUTOTUNE',
    os.path.join(os.path.expanduser('~'), '.cache', 'quantumflow_amd',
                 'autotune.json'))

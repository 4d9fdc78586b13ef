"""Autotuner for the tiled executor and the batched QAOA kernels.

Candidates are timed on the GPU with HIP events; the best configuration per
problem key is stored in a JSON cache (QUANTUMFLOW_AUTOTUNE, default
~/.cache/quantumflow_amd/autotune.json) and reused by ``choose``.  Without
a cached entry ``choose`` returns the built-in defaults below, which are the
configurations measured best on MI355X.
"""

import itertools
import json
import os
import threading

from . import config

_lock = threading.Lock()
_cache = None

# measured best on MI355X for the 28-qubit random-circuit workload
DEFAULTS = {'tile_bits': 12, 'low_bits': 4, 'block': 256, 'fuse_width': 1, 'mode': 'reg',
            'persistent': True}
# A complex128 state of at most 24 qubits (256 MiB) fits MI355X's last-level
# cache: 16-amplitude contiguous runs do not pay there, and 3 low bits leave
# the planner more freedom (100-gate circuits at n = 16, 20, 24: 905 k, 854 k,
# 171 k gates/s against 785 k, 819 k, 143 k with 4 low bits).
SMALL_STATE_QUBITS, SMALL_STATE_LOW_BITS = 24, 3


def _load():
    global _cache
    if _cache is None:
        try:
            with open(config.AUTOTUNE_CACHE) as f:
                _cache = json.load(f)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _save():
    path = config.AUTOTUNE_CACHE
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def block_for(tile_bits, block):
    """Largest allowed block not exceeding the tile."""
    b = block
    while b > 64 and (1 << tile_bits) < b:
        b //= 2 if b > 256 else 4
    return b


def default_config(n):
    cfg = dict(DEFAULTS)
    if config.TILE_BITS:
        cfg['tile_bits'] = config.TILE_BITS
    cfg['tile_bits'] = min(cfg['tile_bits'], n)
    if n <= SMALL_STATE_QUBITS:
        cfg['low_bits'] = SMALL_STATE_LOW_BITS
    cfg['low_bits'] = min(cfg['low_bits'], cfg['tile_bits'])
    cfg['block'] = block_for(cfg['tile_bits'], cfg['block'])
    return cfg


def key_for(kind, n, extra=''):
    return '%s:n=%d%s' % (kind, n, extra)


def choose(n, ngates=0):
    """Configuration for a circuit on n qubits: cached tuning result or the
    defaults."""
    with _lock:
        cached = _load().get(key_for('circuit', n))
    if cached:
        cfg = dict(cached['config'])
        cfg['tile_bits'] = min(cfg['tile_bits'], n)
        cfg['block'] = block_for(cfg['tile_bits'], cfg['block'])
        return cfg
    return default_config(n)


def tune_circuit(gates, qubits, reps=2, candidates=None, store=True, verbose=False):
    """Time candidate plans for this gate list on the GPU and cache the best
    for its qubit count.  Returns (best config, {config tuple: ms})."""
    import torch
    from . import backend as bk
    from .fusion import TilePlan
    n = len(qubits)
    if candidates is None:
        tb = sorted({min(t, n) for t in (11, 12, 13)})
        lb = default_config(n)['low_bits']
        candidates = [dict(tile_bits=t, low_bits=lb, block=b, fuse_width=1, mode='reg')
                      for t, b in itertools.product(tb, (256, 512))]
        candidates += [dict(tile_bits=t, low_bits=lb, block=512, fuse_width=f, mode='lds')
                       for t, f in itertools.product(tb, (1, 3))]
    work = bk.zero_state_tensor(n).reshape(-1).clone()
    results = {}
    best = None
    for cfg in candidates:
        cfg = dict(cfg)
        cfg['block'] = block_for(cfg['tile_bits'], cfg['block'])
        plan = TilePlan(gates, qubits, **cfg)
        plan.run_inplace(work)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            plan.run_inplace(work)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        key = tuple(sorted(cfg.items()))
        results[key] = ms
        if verbose:
            print('autotune', cfg, 'passes', plan.n_passes, 'ms %.3f' % ms, flush=True)
        if best is None or ms < best[1]:
            best = (cfg, ms)
    if store:
        with _lock:
            _load()[key_for('circuit', n)] = {'config': best[0], 'ms': best[1]}
            _save()
    return best[0], results


# measured best on MI355X: 512 ER graphs n 10..22 p=4 (forward) and 64
# 3-regular graphs n 16..24 p=8 (values + adjoint gradients)
QAOA_DEFAULTS = {'forward': {'tile_bits': 13, 'block': 512},
                 'backward': {'tile_bits': 12, 'block': 512}}


def qaoa_key(mode):
    return 'qaoa:' + ('backward' if mode else 'forward')


def choose_qaoa(mode):
    """QAOA-engine configuration {tile_bits, block} for a mode (0 forward,
    1 backward): the cached tuning result, else the built-in default."""
    with _lock:
        cached = _load().get(qaoa_key(mode))
    if cached:
        return dict(cached['config'])
    return dict(QAOA_DEFAULTS['backward' if mode else 'forward'])


def store_qaoa(mode, cfg, ms):
    with _lock:
        _load()[qaoa_key(mode)] = {'config': dict(cfg), 'ms': ms}
        _save()


def tune_qaoa(graphs, p, mode, reps=3, candidates=None, store=True, verbose=False):
    """Time the batched engine on these instances for each candidate
    {tile_bits, block} (one HIP-graph replay per repetition) and cache the
    fastest for the mode.  Returns (best config, {config tuple: ms})."""
    import numpy as np
    import torch
    from . import qaoa_engine as E
    if candidates is None:
        tmax = 12 if mode == E.MODE_BACKWARD else 13
        candidates = [dict(tile_bits=t, block=b)
                      for t, b in itertools.product((tmax, tmax - 1), (256, 512))]
    rng = np.random.default_rng(0)
    beta = rng.uniform(0, np.pi / 2, (len(graphs), p))
    gamma = rng.uniform(0, np.pi, (len(graphs), p))
    results = {}
    best = None
    for cfg in candidates:
        batch = E.QAOABatch(graphs, p, mode=mode, **cfg)
        batch.set_angles(beta, gamma)
        batch.capture()
        batch.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            batch.replay()
        e1.record()
        torch.cuda.synchronize()
        batch.check()
        ms = e0.elapsed_time(e1) / reps
        results[tuple(sorted(cfg.items()))] = ms
        if verbose:
            print('autotune qaoa', cfg, 'ms %.3f' % ms, flush=True)
        if best is None or ms < best[1]:
            best = (dict(cfg), ms)
        del batch
    if store:
        store_qaoa(mode, best[0], best[1])
    return best[0], results

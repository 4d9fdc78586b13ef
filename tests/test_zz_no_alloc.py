"""qfx_expect_diag_noalloc (the expectation whose partials live in the
library's own device array) against qfx_expect_diag with a caller
workspace, through the C-ABI: two calls in a row, the second on a smaller
state, each bit for bit equal (same kernels, same block count), plus the
host-side argument checks."""
import ctypes

import numpy as np
import pytest

from quantumflow_amd import _native

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _expect_both(n, seed):
    import torch
    from quantumflow_amd import backend as bk
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    psi /= np.linalg.norm(psi)
    diag = rng.uniform(-3, 3, size=2 ** n)
    d_psi = torch.from_numpy(psi).to("cuda")
    d_diag = torch.from_numpy(diag).to("cuda")
    lib = _native.load()
    ws = torch.empty(lib.qfx_reduce_workspace_bytes(2 ** n), dtype=torch.uint8, device="cuda")
    r_ws = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    r_na = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    s = bk.stream_handle()
    _native.call('qfx_expect_diag', s, _vp(r_ws), _vp(d_psi), _vp(d_diag), 2 ** n, _vp(ws))
    _native.call('qfx_expect_diag_noalloc', s, _vp(r_na), _vp(d_psi), _vp(d_diag), 2 ** n)
    torch.cuda.synchronize()
    ref = float(np.sum(np.abs(psi) ** 2 * diag))
    return float(r_ws.item()), float(r_na.item()), ref


def test_noalloc_matches_workspace_entry_point_twice():
    import quantumflow_amd  # noqa: F401  (loads the library, picks the device)
    a1, b1, ref1 = _expect_both(20, 1)
    print("n=20: workspace %.17g noalloc %.17g numpy %.17g" % (a1, b1, ref1), flush=True)
    assert a1 == b1
    assert abs(a1 - ref1) <= 1e-12
    # second call on a smaller input: the array keeps the first call's 1024
    # partials, of which this call writes and reads only its own 64
    a2, b2, ref2 = _expect_both(16, 2)
    print("n=16: workspace %.17g noalloc %.17g numpy %.17g" % (a2, b2, ref2), flush=True)
    assert a2 == b2
    assert abs(a2 - ref2) <= 1e-12
    a3, b3, _ = _expect_both(3, 3)
    assert a3 == b3


def test_noalloc_rejects_bad_sizes():
    import torch
    import quantumflow_amd  # noqa: F401
    lib = _native.load()
    x = torch.zeros(4, dtype=torch.complex128, device="cuda")
    d = torch.zeros(4, dtype=torch.float64, device="cuda")
    r = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = lib.qfx_expect_diag_noalloc(None, _vp(r), _vp(x), _vp(d), (1 << 34) + 1)
    assert rc == _native.QFX_ERANGE and 'maximum' in _native.last_error()
    rc = lib.qfx_expect_diag_noalloc(None, _vp(r), _vp(x), _vp(d), 0)
    assert rc == _native.QFX_ERANGE
    rc = lib.qfx_expect_diag_noalloc(None, None, _vp(x), _vp(d), 4)
    assert rc == _native.QFX_EINVAL

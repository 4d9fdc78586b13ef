"""QuantumFlow backend for the AMD MI355X (gfx950): torch-ROCm tensors whose
hot functions run hand-written HIP kernels through the C-ABI in
include/qfx.h (loaded by quantumflow_amd._native).

Exports every name of the reference's backend API
(quantumflow/backend/__init__.py:25-35) with numpy-backend semantics
(quantumflow/backend/numpybk.py, docs/source/backends.rst:71-92):

* tensors are ``torch.complex128`` on the GPU, shaped [2]*N (a view of a
  flat contiguous buffer);
* ``tensormul`` (numpybk.py:159-214), ``inner`` (numpybk.py:125-128) and
  ``absolute`` on complex tensors (used by State.probabilities,
  states.py:113-119) call HIP kernels;
* the extensions ``probabilities``, ``expectation``, ``expectation_z`` and
  ``norm`` give the State readouts (states.py:113-147, qubits.py:180-182)
  as single fused kernels;
* everything else is shape plumbing done with torch on the device, or plain
  numpy on host scalars (gate constructors build their small matrices on
  the host with the same numpy expressions as the reference).

There is no CPU fallback: without libqfx.so or a GPU every hot call raises.
"""

import math
import string
import typing

import numpy as np
import torch

from .. import _native
from ..config import DEVICE, SEED

TL = torch
name = 'rocm'
version = torch.__version__
__version__ = version

CTYPE = torch.complex128
"""Complex dtype of backend tensors"""
FTYPE = torch.float64
"""Real dtype of backend tensors"""
TENSOR = torch.Tensor
BKTensor = typing.Any
TensorLike = typing.Any

MAX_QUBITS = 33
"""Out-of-place gate application holds two states: 2 x 128 GiB at 33 qubits
fits the MI355X's 288 GB of HBM3E."""

EINSUM_SUBSCRIPTS = string.ascii_lowercase + string.ascii_uppercase

pi = np.pi

_NP_CTYPE = np.complex128
_inited = set()
_props = {}
_workspace = {}


# ---------------------------------------------------------------------------
# device plumbing
# ---------------------------------------------------------------------------

def gpu_available() -> bool:
    """True if a GPU is visible to torch and libqfx.so loads."""
    try:
        _native.load()
    except Exception:
        return False
    return torch.cuda.is_available()


def _ensure_device():
    dev = torch.cuda.current_device() if torch.cuda.is_available() else None
    if dev is None:
        raise RuntimeError('quantumflow_amd rocm backend: no GPU visible to torch '
                           '(there is no CPU fallback)')
    if dev not in _inited:
        props = (_native.ctypes.c_int64 * 8)()
        _native.call('qfx_init', props)
        _inited.add(dev)
        _props[dev] = list(props)
    return dev


def device_props():
    """qfx_init's properties of the current device: CUs, LDS bytes per
    block, warp size, L2 bytes, clock kHz, total MiB, max threads per block,
    wall-clock rate kHz."""
    return list(_props[_ensure_device()])


def stream_handle():
    """hipStream_t of torch's current stream (0 = default stream)."""
    _ensure_device()
    return torch.cuda.current_stream().cuda_stream


def workspace(nbytes):
    """Device scratch of at least nbytes for reductions, reused in stream
    order (all qfx calls go to torch's current stream)."""
    dev = _ensure_device()
    key = (dev, torch.cuda.current_stream().cuda_stream)
    ws = _workspace.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(int(nbytes), 1 << 20), dtype=torch.uint8,
                         device='cuda')
        _workspace[key] = ws
    return ws


def _ptr(t):
    return _native.ctypes.c_void_p(t.data_ptr())


def as_device(tensor):
    """Flat contiguous complex128 device tensor (no copy if already one)."""
    _ensure_device()
    if isinstance(tensor, torch.Tensor):
        t = tensor
        if t.dtype != CTYPE or t.device.type != 'cuda':
            t = t.to(device=DEVICE, dtype=CTYPE)
    else:
        t = torch.from_numpy(np.ascontiguousarray(
            np.asarray(tensor, dtype=_NP_CTYPE))).to(DEVICE)
    if not t.is_contiguous():
        t = t.contiguous()
    return t.reshape(-1)


def _host(tensor):
    if isinstance(tensor, torch.Tensor):
        return tensor.detach().cpu().numpy()
    return np.asarray(tensor)


# ---------------------------------------------------------------------------
# conversions (numpybk.py:79-112)
# ---------------------------------------------------------------------------

def ccast(value: complex) -> TensorLike:
    """Cast value to complex tensor (if necessary)"""
    return value


def fcast(value: float) -> TensorLike:
    """Cast value to float tensor (if necessary)"""
    return value


# Host arrays of gate size (2^k x 2^k, k <= 5) are classified when they are
# uploaded, and the class rides on the device tensor (with the tensor's
# version counter, so an in-place change forgets it): the reference's
# Gate.run passes such a device tensor to tensormul for every application
# (quantumflow/ops.py:161-166), which then needs no device-to-host copy.
_KIND_MAX_ELEMS = 4 ** 5


def _remember_kind(t, kind):
    t._qfx_kind = (t._version, kind)
    return t


def _known_kind(t):
    rec = getattr(t, '_qfx_kind', None)
    if rec is not None and rec[0] == t._version:
        return rec[1]
    return None


def _gate_kind_of_host(arr):
    size = arr.size
    if 4 <= size <= _KIND_MAX_ELEMS and size & (size - 1) == 0 and \
            int(math.log2(size)) % 2 == 0:
        dim = int(round(math.sqrt(size)))
        return GateKind(arr.reshape(dim, dim))
    return None


def astensor(array: TensorLike) -> BKTensor:
    """Convert to a complex128 device tensor (shape preserved)."""
    _ensure_device()
    if isinstance(array, torch.Tensor):
        t = array.to(device=DEVICE, dtype=CTYPE)
        if t is not array or array.device.type != 'cuda':
            kind = _known_kind(array) if array.device.type == 'cuda' else None
            if kind is None and array.device.type != 'cuda':
                kind = _gate_kind_of_host(array.detach().numpy().astype(_NP_CTYPE))
            if kind is not None:
                _remember_kind(t, kind)
        return t
    arr = np.ascontiguousarray(np.asarray(array, dtype=_NP_CTYPE))
    t = torch.from_numpy(arr).to(DEVICE)
    kind = _gate_kind_of_host(arr)
    return _remember_kind(t, kind) if kind is not None else t


def astensorproduct(array: TensorLike) -> BKTensor:
    """astensor, reshaped to [2]*N (numpybk.py:105-112)."""
    tensor = astensor(array)
    n = int(math.log2(tensor.numel()))
    out = tensor.reshape([2] * n)
    kind = _known_kind(tensor)
    return _remember_kind(out, kind) if kind is not None and out is not tensor else out


def evaluate(tensor: BKTensor) -> TensorLike:
    """The value as a numpy object (device -> host copy; the only sync)."""
    if isinstance(tensor, torch.Tensor):
        arr = tensor.detach().cpu().numpy()
        return arr[()] if arr.ndim == 0 else arr
    return tensor


def rank(tensor: BKTensor) -> int:
    return len(tensor.shape)


def size(tensor: BKTensor) -> int:
    if isinstance(tensor, torch.Tensor):
        return tensor.numel()
    return np.size(tensor)


def set_random_seed(seed: int) -> None:
    np.random.seed(seed)
    torch.manual_seed(seed)


def getitem(tensor: BKTensor, key: typing.Any) -> BKTensor:
    return tensor.__getitem__(key)


# ---------------------------------------------------------------------------
# elementwise / shape functions: torch on device tensors, numpy otherwise
# ---------------------------------------------------------------------------

def _dispatch(tfn, nfn):
    def f(x, *args, **kwargs):
        if isinstance(x, torch.Tensor):
            return tfn(x, *args, **kwargs)
        return nfn(x, *args, **kwargs)
    f.__name__ = nfn.__name__
    return f


conj = _dispatch(torch.conj, np.conj)
cos = _dispatch(torch.cos, np.cos)
sin = _dispatch(torch.sin, np.sin)
exp = _dispatch(torch.exp, np.exp)
sqrt = _dispatch(torch.sqrt, np.sqrt)
arccos = _dispatch(torch.arccos, np.arccos)
imag = _dispatch(lambda t: torch.imag(t) if t.is_complex() else torch.zeros_like(t), np.imag)
diag = _dispatch(torch.diag, np.diag)
trace = _dispatch(torch.trace, np.trace)


def real(tensor):
    if isinstance(tensor, torch.Tensor):
        return torch.real(tensor) if tensor.is_complex() else tensor
    return np.real(tensor)


def minimum(a, b):
    if isinstance(a, torch.Tensor) or isinstance(b, torch.Tensor):
        a = a if isinstance(a, torch.Tensor) else torch.tensor(a, device=DEVICE)
        b = b if isinstance(b, torch.Tensor) else torch.tensor(b, device=a.device,
                                                                dtype=a.dtype)
        return torch.minimum(a, b)
    return np.minimum(a, b)


def sum(tensor, *args, **kwargs):  # noqa: A001 (numpy name)
    if isinstance(tensor, torch.Tensor):
        return torch.sum(tensor, *args, **kwargs)
    return np.sum(tensor, *args, **kwargs)


def reshape(tensor, shape):
    if isinstance(tensor, torch.Tensor):
        return torch.reshape(tensor, tuple(shape))
    return np.reshape(tensor, shape)


def transpose(tensor, perm=None):
    if isinstance(tensor, torch.Tensor):
        if perm is None:
            perm = list(reversed(range(tensor.dim())))
        return tensor.permute(*perm).contiguous()
    return np.transpose(tensor, perm)


def matmul(a, b):
    if isinstance(a, torch.Tensor) or isinstance(b, torch.Tensor):
        return torch.matmul(astensor(a), astensor(b))
    return np.matmul(a, b)


def outer(a, b):
    if isinstance(a, torch.Tensor) or isinstance(b, torch.Tensor):
        a = astensor(a).reshape(-1)
        b = astensor(b).reshape(-1)
        return torch.outer(a, b)
    return np.outer(a, b)


def einsum(subscripts, *operands):
    if any(isinstance(o, torch.Tensor) for o in operands):
        return torch.einsum(subscripts, *[astensor(o) for o in operands])
    return np.einsum(subscripts, *operands)


def cis(theta: float) -> BKTensor:
    """cis(theta) = exp(i theta) (numpybk.py:131-137)."""
    if isinstance(theta, torch.Tensor):
        return torch.exp(1j * theta)
    return np.exp(theta * 1.0j)


def productdiag(tensor: BKTensor) -> BKTensor:
    """Diagonal of the product tensor as [2]*(N/2) (numpybk.py:150-156)."""
    n = rank(tensor)
    t = reshape(tensor, [2 ** (n // 2), 2 ** (n // 2)])
    t = diag(t)
    return reshape(t, [2] * (n // 2))


# ---------------------------------------------------------------------------
# the hot path: gate application
# ---------------------------------------------------------------------------

class GateKind:
    """Classification of a 2^k x 2^k gate matrix for kernel selection."""
    DENSE, DIAGONAL, MONOMIAL = 'dense', 'diagonal', 'monomial'

    __slots__ = ('kind', 'k', 'matrix', 'diag', 'src', 'phase')

    def __init__(self, matrix):
        m = np.asarray(matrix, dtype=_NP_CTYPE)
        dim = int(round(math.sqrt(m.size)))
        m = m.reshape(dim, dim)
        self.k = int(math.log2(dim))
        self.matrix = np.ascontiguousarray(m)
        self.diag = self.src = self.phase = None
        nz = m != 0
        if self.k <= 5 and not np.any(nz & ~np.eye(dim, dtype=bool)):
            self.kind = GateKind.DIAGONAL
            self.diag = np.diag(m).copy()
        elif self.k <= 5 and np.all(nz.sum(axis=1) == 1) and np.all(nz.sum(axis=0) == 1):
            self.kind = GateKind.MONOMIAL
            self.src = np.argmax(nz, axis=1).astype(np.int32)
            ph = m[np.arange(dim), self.src]
            self.phase = None if np.all(ph == 1.0) else ph
        else:
            self.kind = GateKind.DENSE


def classify(matrix):
    return GateKind(matrix)


def apply_gate(out, state, n, bits, kind, inplace_sparse=False):
    """Launch the kernel for `kind` on flat device tensors.  ``bits`` are
    index bit positions (n-1-axis) in gate-qubit order."""
    s = stream_handle()
    k = kind.k
    tb = _native.int_array(bits)
    if kind.kind == GateKind.DIAGONAL:
        d = _native.c128_array(kind.diag)
        if inplace_sparse and out.data_ptr() == state.data_ptr():
            _native.call('qfx_apply_diagonal_inplace', s, _ptr(out), n, k, tb, d)
        else:
            _native.call('qfx_apply_diagonal', s, _ptr(out), _ptr(state), n, k, tb, d)
    elif kind.kind == GateKind.MONOMIAL:
        src = _native.int_array(kind.src)
        ph = None if kind.phase is None else _native.c128_array(kind.phase)
        if inplace_sparse and out.data_ptr() == state.data_ptr():
            _native.call('qfx_apply_monomial_inplace', s, _ptr(out), n, k, tb, src, ph)
        else:
            _native.call('qfx_apply_monomial', s, _ptr(out), _ptr(state), n, k, tb, src, ph)
    elif k <= 3:
        m = _native.c128_array(kind.matrix.reshape(-1))
        _native.call('qfx_apply_matrix', s, _ptr(out), _ptr(state), n, k, tb, m)
    else:
        dm = torch.from_numpy(kind.matrix.reshape(-1)).to(out.device)
        if out.data_ptr() == state.data_ptr():
            tmp = torch.empty_like(state)
            _native.call('qfx_apply_matrix_dev', s, _ptr(tmp), _ptr(state), n, k, tb, _ptr(dm))
            out.copy_(tmp)
        else:
            _native.call('qfx_apply_matrix_dev', s, _ptr(out), _ptr(state), n, k, tb, _ptr(dm))
    return out


def tensormul(tensor0: BKTensor, tensor1: BKTensor,
              indices: typing.List[int]) -> BKTensor:
    """Contract gate ``tensor0`` (rank 2K, [ket_out..., bra_in...]) into
    ``tensor1`` (rank N) on axes ``indices``; returns a new [2]*N device
    tensor (numpybk.py:159-214).  ``tensor0`` may be a GateKind, a host
    array (classified on the host) or a device tensor.  A device tensor made
    by ``astensor``/``astensorproduct`` from a host array carries its class;
    any other device tensor is applied as a dense matrix read from device
    memory.  Neither copies anything to the host."""
    dev_matrix = None
    if isinstance(tensor0, GateKind):
        kind = tensor0
        k = kind.k
    else:
        k = rank(tensor0) // 2
        kind = None
    n = rank(tensor1)
    assert k == len(indices)
    if kind is None:
        if isinstance(tensor0, torch.Tensor) and tensor0.device.type == 'cuda':
            kind = _known_kind(tensor0)
            if kind is None:
                if k > 12:
                    raise ValueError('tensormul: gates of more than 12 qubits are not supported')
                dev_matrix = as_device(tensor0)
        else:
            kind = classify(_host(tensor0).reshape(2 ** k, 2 ** k))
    state = as_device(tensor1)
    if n != int(math.log2(state.numel())):
        raise ValueError('tensormul: state is not a product tensor')
    out = torch.empty_like(state)
    bits = [n - 1 - int(i) for i in indices]
    if dev_matrix is not None:
        _native.call('qfx_apply_matrix_dev', stream_handle(), _ptr(out), _ptr(state), n, k,
                     _native.int_array(bits), _ptr(dev_matrix))
    else:
        apply_gate(out, state, n, bits, kind)
    return out.view([2] * n)


# ---------------------------------------------------------------------------
# readouts
# ---------------------------------------------------------------------------

def inner(tensor0: BKTensor, tensor1: BKTensor) -> BKTensor:
    """sum conj(t0) t1 over the flattened tensors (np.vdot), as a 0-dim
    complex128 device tensor (numpybk.py:125-128)."""
    a = as_device(tensor0)
    b = as_device(tensor1)
    if a.numel() != b.numel():
        raise ValueError('inner: tensors differ in size')
    res = torch.empty((), dtype=CTYPE, device=a.device)
    ws = workspace(_native.load().qfx_reduce_workspace_bytes(a.numel()))
    _native.call('qfx_inner', stream_handle(), _ptr(res), _ptr(a), _ptr(b), a.numel(), _ptr(ws))
    return res


def absolute(tensor: BKTensor) -> BKTensor:
    """|x| elementwise; complex device tensors use the hypot kernel."""
    if isinstance(tensor, torch.Tensor):
        if tensor.is_complex() and tensor.dim() > 0:
            a = as_device(tensor)
            out = torch.empty(a.numel(), dtype=FTYPE, device=a.device)
            _native.call('qfx_abs', stream_handle(), _ptr(out), _ptr(a), a.numel())
            return out.view(tensor.shape)
        return torch.abs(tensor)
    return np.absolute(tensor)


def probabilities(tensor: BKTensor) -> BKTensor:
    """|psi|^2 as float64 [2]*N in one kernel (states.py:113-119)."""
    a = as_device(tensor)
    out = torch.empty(a.numel(), dtype=FTYPE, device=a.device)
    _native.call('qfx_abs2', stream_handle(), _ptr(out), _ptr(a), a.numel())
    return out.view(tuple(tensor.shape))


def _real_diag(diag, numel):
    if isinstance(diag, torch.Tensor):
        d = diag.to(device=DEVICE)
        d = torch.real(d) if d.is_complex() else d
        d = d.to(FTYPE)
    else:
        d = torch.from_numpy(np.ascontiguousarray(
            np.real(np.asarray(diag, dtype=_NP_CTYPE)).astype(np.float64))).to(DEVICE)
    d = d.contiguous().reshape(-1)
    if d.numel() != numel:
        raise ValueError('expectation: diagonal does not match the state size')
    return d


def expectation(diag_hermitian: BKTensor, tensor: BKTensor) -> BKTensor:
    """sum real(diag) |psi|^2 as a 0-dim float64 device tensor, one fused
    reduction (states.py:131-147, exact mode)."""
    a = as_device(tensor)
    d = _real_diag(diag_hermitian, a.numel())
    res = torch.empty((), dtype=FTYPE, device=a.device)
    ws = workspace(_native.load().qfx_reduce_workspace_bytes(a.numel()))
    _native.call('qfx_expect_diag', stream_handle(), _ptr(res), _ptr(a), _ptr(d), a.numel(),
                 _ptr(ws))
    return res


def expectation_z(tensor: BKTensor) -> BKTensor:
    """<Z_a> for every axis a in one pass; float64 device tensor [N]."""
    a = as_device(tensor)
    n = int(math.log2(a.numel()))
    out = torch.empty(n, dtype=FTYPE, device=a.device)
    ws = workspace(_native.load().qfx_reduce_workspace_bytes(a.numel()))
    _native.call('qfx_expect_z', stream_handle(), _ptr(out), _ptr(a), n, _ptr(ws))
    return out


def qubit_probs(tensor: BKTensor, axis: int) -> BKTensor:
    """[P(axis in |0>), P(axis in |1>)] (unnormalised sums of |psi|^2), float64
    device tensor [2], one reduction pass (the P0(q).run(ket).norm() of
    stdops.py:53-54 without forming the projected state)."""
    a = as_device(tensor)
    n = int(math.log2(a.numel()))
    axis = int(axis)
    if not 0 <= axis < n:
        raise ValueError('qubit_probs: axis out of range')
    out = torch.empty(2, dtype=FTYPE, device=a.device)
    ws = workspace(_native.load().qfx_reduce_workspace_bytes(a.numel()))
    _native.call('qfx_qubit_probs', stream_handle(), _ptr(out), _ptr(a), n, axis, _ptr(ws))
    return out


def norm(tensor: BKTensor) -> BKTensor:
    """|<psi|psi>| (the squared 2-norm, qubits.py:180-182), 0-dim float64."""
    a = as_device(tensor)
    res = torch.empty((), dtype=FTYPE, device=a.device)
    ws = workspace(_native.load().qfx_reduce_workspace_bytes(a.numel()))
    _native.call('qfx_norm2', stream_handle(), _ptr(res), _ptr(a), a.numel(), _ptr(ws))
    return res


def draw_seed() -> int:
    """A 64-bit seed for the device sampler, drawn from numpy's global
    generator, so np.random.seed / set_random_seed make device sampling
    reproducible as they do the reference's numpy draws."""
    return int(np.random.randint(0, 2 ** 63 - 1, dtype=np.int64))


def sample_indices(tensor: BKTensor, trials: int, seed: int) -> BKTensor:
    """trials basis indices drawn from |psi|^2 (normalised on the device),
    int64 device tensor; index bit N-1-a is qubit axis a."""
    a = as_device(tensor)
    n = int(math.log2(a.numel()))
    if 2 ** n != a.numel() or n < 1:
        raise ValueError('sample: state size is not a power of two')
    trials = int(trials)
    if trials < 0:
        raise ValueError('sample: negative number of trials')
    out = torch.empty(trials, dtype=torch.int64, device=a.device)
    if trials == 0:
        return out
    ws = workspace(_native.load().qfx_sample_workspace_bytes(n))
    _native.call('qfx_sample', stream_handle(), _ptr(out), _ptr(a), n, trials,
                 int(seed) & (2 ** 64 - 1), _ptr(ws))
    return out


def sample_counts(tensor: BKTensor, trials: int, seed: int) -> BKTensor:
    """Counts of each basis index over trials device draws, int64 [2]*N
    (the shape of np.random.multinomial(trials, probs).reshape([2]*N))."""
    a = as_device(tensor)
    n = int(math.log2(a.numel()))
    idx = sample_indices(a, trials, seed)
    counts = torch.zeros(a.numel(), dtype=torch.int64, device=a.device)
    _native.call('qfx_histogram', stream_handle(), _ptr(counts), _ptr(idx), idx.numel(), n)
    return counts.view([2] * n)


def sample_expectation(diag_hermitian: BKTensor, tensor: BKTensor, trials: int,
                       seed: int) -> BKTensor:
    """sum_i real(diag_i) counts_i / trials over ``trials`` device draws
    (states.py:140-147 with trials given), 0-dim float64 device tensor."""
    a = as_device(tensor)
    d = _real_diag(diag_hermitian, a.numel())
    idx = sample_indices(a, trials, seed)
    return torch.sum(d[idx]) / int(trials)


def zero_state_tensor(n: int) -> BKTensor:
    """|0...0> on n qubits, filled on the device."""
    _ensure_device()
    t = torch.empty(2 ** n, dtype=CTYPE, device=DEVICE)
    _native.call('qfx_zero_state', stream_handle(), _ptr(t), n)
    return t.view([2] * n)


if SEED is not None:
    set_random_seed(SEED)

"""Batched QAOA MaxCut on the GPU: values, adjoint gradients, optimizer.

The reference evaluates one graph at a time: qubo_circuit
(quantumflow/qaoa.py:22-64) -> Circuit.run -> State.expectation(graph_cuts)
(quantumflow/states.py:131-147, quantumflow/qaoa.py:68-81), and gets
gradients from TensorFlow autograd, which stores every intermediate state
(examples/qaoa_maxcut.py:53-64).  ``QAOABatch`` evaluates many graphs of
different sizes in one persistent kernel launch (csrc/qaoa_engine.hip), and
in QFX_QAOA_BACKWARD mode also returns d<C>/d(beta, gamma) by the adjoint
method, keeping two states (psi and lambda) per instance instead of one per
gate.  ``QAOAExpectation`` puts it behind torch.autograd; ``optimize`` runs
the reference's plain gradient ascent for all instances at once.

Program of one instance (built on the host by ``build_program``): the gate
sequence H^n, [cost(l), RX(l) on every qubit] for l < p, then <C>, and in
backward mode lambda = C psi followed by the stages undone in reverse.  It is
split into passes over tile bit sets of at most T bits (13 forward, 12
backward): the low 13 (12) bits, then the low 3 bits plus the next chunk of
high bits, and so on; an RX stage on bits outside the current set closes the
pass.  Instances with n <= T run as one pass entirely in LDS.
"""

import ctypes
import math

import networkx as nx
import numpy as np
import torch

from . import _native
from . import backend as bk

MODE_FORWARD, MODE_BACKWARD = 0, 1
ST_INIT, ST_COST, ST_RX, ST_EXPECT, ST_MAKE_LAMBDA, ST_RX_ADJ, ST_COST_ADJ = 1, 2, 3, 4, 5, 6, 7
# fused stage: cost layer l then RX layer l (the phases ride in the first
# mixer chunk)
ST_COST_RX = 8
# RX layer l on tile bits applied in registers while the tile is loaded
# from / stored to HBM (the pass's first / last stage)
ST_RX_LOAD, ST_RX_STORE = 10, 11
IO_BITS = 4               # tile bits above log2(block) a thread holds in registers
PF_LOAD_PSI, PF_LOAD_LAM, PF_STORE_PSI, PF_STORE_LAM = 1, 2, 4, 8
MAX_LAYERS = 16
MAX_STAGES = 40
LOW_BITS = 3


class QaoaInst(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int32), ('p', ctypes.c_int32), ('weighted', ctypes.c_int32),
                ('nedges', ctypes.c_int32), ('wsum', ctypes.c_double),
                ('cut_off', ctypes.c_int64), ('psi_off', ctypes.c_int64),
                ('lam_off', ctypes.c_int64), ('angle_off', ctypes.c_int32),
                ('slot0', ctypes.c_int32), ('nslots', ctypes.c_int32), ('mirror', ctypes.c_int32)]


class QaoaPass(ctypes.Structure):
    _fields_ = [('T', ctypes.c_int32), ('pos', ctypes.c_int32 * 13), ('nstage', ctypes.c_int32),
                ('flags', ctypes.c_int32), ('stage', (ctypes.c_int32 * 3) * MAX_STAGES)]


class QaoaItem(ctypes.Structure):
    _fields_ = [('inst', ctypes.c_int32), ('pass_', ctypes.c_int32), ('tile', ctypes.c_int32),
                ('partial', ctypes.c_int32)]


class QaoaArgs(ctypes.Structure):
    _fields_ = [(nm, ctypes.c_void_p) for nm in
                ('insts', 'passes', 'items', 'kstart', 'angles', 'psi_pool', 'lam_pool',
                 'cnt_pool', 'val_pool', 'partials', 'out', 'sync')] + \
               [(nm, ctypes.c_int32) for nm in
                ('ninst', 'nk', 'width', 'mode', 'block', 'max_items_per_k', 'tile_bits',
                 'pad')] + \
               [('timing', ctypes.c_void_p), ('timing_len', ctypes.c_int64)]


def tile_sets(n, T, low=None):
    """Tile bit sets covering n index bits: the lowest T bits, then the low
    `low` bits plus successive chunks of T - low higher bits, each padded to
    T bits with the lowest bits it does not hold (bigger tiles, fewer work
    items, and more bits on which pending RX stages can run).  By default
    `low` is 3 (128-byte contiguous runs), lowered to 2 or 1 when that is
    what lets two sets cover all bits (two passes per layer instead of four
    outweigh the shorter runs of the second set)."""
    if n <= T:
        return [tuple(range(n))]
    if low is None:
        low = LOW_BITS if 2 * T - n >= LOW_BITS else (2 * T - n if 2 * T - n >= 1 else LOW_BITS)
    low = min(low, T - 1)               # every later set must take a new bit
    sets = [tuple(range(T))]
    lo = list(range(low))
    b = T
    while b < n:
        chunk = list(range(b, min(n, b + T - low)))
        st = set(lo + chunk)
        j = 0
        while len(st) < T:
            st.add(j)
            j += 1
        sets.append(tuple(sorted(st)))
        b += T - low
    return sets


def tile_sets_mirror(n, T, low=None):
    """Tile bit sets for a mirror instance: like tile_sets over the bits
    0..n-2 with T-1 bits each, every set then completed by bit n-1 as its
    highest tile bit."""
    top = n - 1
    return [tuple(sorted(set(s_) | {top})) for s_ in tile_sets(n - 1, T - 1, low)]


def use_mirror(n, T):
    """Run an n-qubit instance on half its state (complement symmetry): when
    it needs more than one pass and leaves at least one non-tile bit besides
    the mirrored one (n >= T + 2)."""
    return n >= T + 2 and T >= 2


def fuse_stages(stages):
    """Merge a cost stage into the mixer stage of the same layer that follows
    it in a pass: (COST l, RX l) -> COST_RX l.  The kernel then applies the
    phases inside the first RX chunk instead of in a sweep of their own.
    The adjoint stages stay separate: with lambda and psi both in registers,
    a fused adjoint chunk measured slower than the two sweeps."""
    out = []
    for st in stages:
        typ, layer, bits = st
        if out:
            pt, pl, _ = out[-1]
            if pt == ST_COST and typ == ST_RX and pl == layer and bits:
                out[-1] = (ST_COST_RX, layer, bits)
                continue
        out.append(st)
    return out


def io_stages(tset, stages, flags, block):
    """Move RX bits that a thread holds in registers during the tile's HBM
    load or store out of the pass's first or last mixer stage.  Thread t
    owns tile elements t + k*block and holds 2^IO_BITS consecutive k at a
    time, so the tile bits log2(block) .. +IO_BITS-1 pair elements of one
    thread: RX on them needs no LDS sweep.  The first stage gives up such bits to ST_RX_LOAD when the
    pass loads the state; the last to ST_RX_STORE when it stores it.  A
    stage keeps at least one bit (a fused cost stage needs its chunk)."""
    lb = block.bit_length() - 1
    kset = {b for j, b in enumerate(tset) if lb <= j < lb + IO_BITS}
    stages = list(stages)
    if (flags & PF_LOAD_PSI) and stages and stages[0][0] == ST_RX:
        typ, layer, bits = stages[0]
        mv = frozenset(bits & kset)
        if mv:
            rest = frozenset(bits - mv)
            stages[0:1] = [(ST_RX_LOAD, layer, mv)] + ([(typ, layer, rest)] if rest else [])
    if (flags & PF_STORE_PSI) and stages and stages[-1][0] in (ST_RX, ST_COST_RX):
        typ, layer, bits = stages[-1]
        mv = frozenset(bits & kset)
        if mv and bits - mv:
            stages[-1:] = [(typ, layer, frozenset(bits - mv)), (ST_RX_STORE, layer, mv)]
    return stages


def build_program(n, p, mode, T, fuse=True, block=512, mirror=False):
    """Passes of one instance: list of (tile bits, [(stage, layer, global bit
    set)], flags).  With `fuse`, forward-mode programs merge each cost stage
    into the mixer stage of its layer (fuse_stages) and apply RX bits held
    in registers during the tile's load and store (io_stages, for
    `block`-thread workgroups).  Backward programs keep their cost stages
    separate: their 12-bit tiles leave a 4-bit chunk only 256 groups, half
    of a 512-thread workgroup, and the fused stage measured slower there
    than the separate sweep."""
    allbits = frozenset(range(n))
    seq = [(ST_INIT, 0, None)]
    for l in range(p):
        seq.append((ST_COST, l, None))
        seq.append((ST_RX, l, allbits))
    seq.append((ST_EXPECT, 0, None))
    if mode == MODE_BACKWARD:
        seq.append((ST_MAKE_LAMBDA, 0, None))
        for l in reversed(range(p)):
            seq.append((ST_RX_ADJ, l, allbits))
            seq.append((ST_COST_ADJ, l, None))
    sets = tile_sets_mirror(n, T) if mirror else tile_sets(n, T)
    passes = []
    cur = [sets[0], []]
    for typ, layer, bits in seq:
        if bits is None:
            cur[1].append((typ, layer, None))
            continue
        remaining = set(bits)
        while remaining:
            here = remaining & set(cur[0])
            if here:
                cur[1].append((typ, layer, frozenset(here)))
                remaining -= here
            if remaining:
                passes.append(cur)
                best = max(sets, key=lambda s_: len(remaining & set(s_)))
                cur = [best, []]
    passes.append(cur)
    # load/store flags
    out = []
    lam_live = False
    for q, (tset, stages) in enumerate(passes):
        flags = 0
        if q > 0:
            flags |= PF_LOAD_PSI
            if lam_live:
                flags |= PF_LOAD_LAM
        if any(st[0] == ST_MAKE_LAMBDA for st in stages):
            lam_live = True
        if q < len(passes) - 1:
            flags |= PF_STORE_PSI
            if lam_live:
                flags |= PF_STORE_LAM
        if fuse and mode == MODE_FORWARD:
            stages = io_stages(tuple(tset), fuse_stages(stages), flags, block)
        out.append((tuple(tset), stages, flags))
    return out


def program_bytes(n, p, mode, T, block=512, mirror=False, weighted=False, fuse=True):
    """Algorithmic HBM bytes of one instance's program: (state, cut table).
    State: 16 B per kept amplitude (2^n, or 2^(n-1) for a mirror instance)
    for every load and store of psi and lambda a pass makes (a one-pass
    instance keeps its state in LDS).  Cut table: 2 B (unit weights) or 8 B
    per kept amplitude for every stage that reads the cost of each index
    (cost layers, fused cost-mixer stages, <C>, lambda = C psi, adjoint cost
    layers)."""
    prog = build_program(n, p, mode, T, fuse, block, mirror=mirror)
    kept = 2 ** (n - 1) if mirror else 2 ** n
    reads = (ST_COST, ST_COST_RX, ST_EXPECT, ST_MAKE_LAMBDA, ST_COST_ADJ)
    state = cut = 0
    for _, stages, flags in prog:
        if len(prog) > 1:
            state += 16 * kept * sum(bool(flags & f) for f in
                                     (PF_LOAD_PSI, PF_LOAD_LAM, PF_STORE_PSI, PF_STORE_LAM))
        cut += (8 if weighted else 2) * kept * sum(st[0] in reads for st in stages)
    return state, cut


def graph_arrays(graph):
    n = len(graph)
    if sorted(graph.nodes()) != list(range(n)):
        raise ValueError('QAOA graphs need node labels 0..n-1')
    edges = [(u, v) for u, v in graph.edges()]
    ws = [graph[u][v].get('weight', None) for u, v in edges]
    weighted = any(w is not None and float(w) != 1.0 for w in ws)
    w = np.array([1.0 if x is None else float(x) for x in ws], dtype=np.float64)
    return n, np.array(edges, dtype=np.int32).reshape(-1, 2), w, weighted


class QAOABatch:
    """A fixed batch of MaxCut instances compiled for the GPU engine.

    graphs: networkx graphs with nodes 0..n-1 (edge attribute 'weight'
    optional); p: layers; mode: MODE_FORWARD (values) or MODE_BACKWARD
    (values and gradients)."""

    def __init__(self, graphs, p, mode=MODE_BACKWARD, block=None, tile_bits=None, fuse=True,
                 mirror=True):
        if not 1 <= p <= MAX_LAYERS:
            raise ValueError('p must be 1..%d' % MAX_LAYERS)
        if block is None or tile_bits is None:
            from .autotune import choose_qaoa
            tuned = choose_qaoa(mode)
            block = tuned['block'] if block is None else block
            tile_bits = tuned['tile_bits'] if tile_bits is None else tile_bits
        self.graphs = list(graphs)
        self.p = p
        self.mode = mode
        self._mirror = bool(mirror)
        self._fuse = fuse
        self.block = block
        self.width = 1 if mode == MODE_FORWARD else 1 + 2 * p
        T_max = 13 if mode == MODE_FORWARD else 12
        T = T_max if tile_bits is None else int(tile_bits)
        if not 6 <= T <= T_max:
            raise ValueError('tile_bits must be 6..%d in this mode' % T_max)
        self.T = T
        dev = bk.DEVICE
        bk.stream_handle()                          # initialise the device
        ninst = len(self.graphs)
        insts = (QaoaInst * ninst)()
        pass_list, items_by_k = [], []
        psi_len = lam_len = 0
        cnt_len = val_len = 0
        slot = 0
        single_max = 1
        self.n = []
        self.maxcut = np.zeros(ninst)
        cut_jobs = []
        for i, g in enumerate(self.graphs):
            n, edges, w, weighted = graph_arrays(g)
            self.n.append(n)
            mir = bool(mirror) and use_mirror(n, T)
            prog = build_program(n, p, mode, T, fuse, block, mirror=mir)
            inst = insts[i]
            inst.n, inst.p, inst.weighted, inst.nedges = n, p, int(weighted), len(edges)
            inst.mirror = int(mir)
            kept = 2 ** (n - 1) if mir else 2 ** n          # amplitudes in the pools
            inst.wsum = float(np.sum(w))
            if weighted:
                inst.cut_off = val_len
                val_len += 2 ** n
            else:
                inst.cut_off = cnt_len
                cnt_len += 2 ** n
            cut_jobs.append((n, edges, w, weighted, inst.cut_off))
            if len(prog) > 1:
                inst.psi_off = psi_len
                psi_len += kept
                if mode == MODE_BACKWARD:
                    inst.lam_off = lam_len
                    lam_len += kept
            else:
                # a one-pass instance never touches the pools; its offsets
                # point at the start, which the pools are sized to cover
                single_max = max(single_max, 2 ** n)
            inst.angle_off = 2 * p * i
            inst.slot0 = slot
            for q, (tset, stages, flags) in enumerate(prog):
                ps = QaoaPass()
                ps.T = len(tset)
                for j, b in enumerate(tset):
                    ps.pos[j] = b
                ps.nstage = len(stages)
                if len(stages) > MAX_STAGES:
                    raise ValueError('program pass has too many stages (p too large)')
                local = {b: j for j, b in enumerate(tset)}
                accum = False
                for si, (typ, layer, bits) in enumerate(stages):
                    mask = 0
                    if bits is not None:
                        for b in bits:
                            mask |= 1 << local[b]
                    ps.stage[si][0], ps.stage[si][1], ps.stage[si][2] = typ, layer, mask
                    accum |= typ in (ST_EXPECT, ST_RX_ADJ, ST_COST_ADJ)
                ps.flags = flags
                pidx = len(pass_list)
                pass_list.append(ps)
                while len(items_by_k) <= q:
                    items_by_k.append([])
                # a mirror instance's items: tiles whose highest non-tile bit is clear
                for t in range(2 ** (n - ps.T - (1 if mir else 0))):
                    items_by_k[q].append((i, pidx, t, slot if accum else -1, 2 ** n))
                    if accum:
                        slot += 1
            inst.nslots = slot - inst.slot0
        self.dummy_slot = slot            # items without a reduction write here
        nslots = slot + 1
        # global pass k: single-pass (whole-instance) items first, then big tiles
        items, kstart = [], [0]
        for k, lst in enumerate(items_by_k):
            lst.sort(key=lambda x: (-(pass_list[x[1]].T == self.n[x[0]]) * x[4], x[0], x[2]))
            for (i, pidx, t, sl, _) in lst:
                it = QaoaItem()
                it.inst, it.pass_, it.tile = i, pidx, t
                it.partial = sl if sl >= 0 else self.dummy_slot
                items.append(it)
            kstart.append(len(items))
        self.nk = len(items_by_k)
        self.nitems = len(items)
        psi_len = max(psi_len, single_max)
        lam_len = max(lam_len, single_max)
        # host-side validation of the whole program
        passes_arr = (QaoaPass * len(pass_list))(*pass_list)
        items_arr = (QaoaItem * len(items))(*items)
        ks = (ctypes.c_int * len(kstart))(*kstart)
        lib = _native.load()
        _native.check(lib.qfx_qaoa_check(insts, ninst, passes_arr, len(pass_list), items_arr, ks,
                                         self.nk, self.width, psi_len,
                                         lam_len if mode == MODE_BACKWARD else 0,
                                         max(cnt_len, 1) if cnt_len else 0,
                                         max(val_len, 1) if val_len else 0, nslots, mode, T),
                      'qfx_qaoa_check')

        def up(arr):
            return torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
        self._insts = up(insts)
        self._passes = up(passes_arr)
        self._items = up(items_arr)
        self._kstart = torch.tensor(kstart, dtype=torch.int32, device=dev)
        self.psi_pool = torch.empty(psi_len, dtype=torch.complex128, device=dev)
        self.lam_pool = torch.empty(lam_len if mode == MODE_BACKWARD else 1,
                                    dtype=torch.complex128, device=dev)
        self.cnt_pool = torch.zeros(max(cnt_len, 1), dtype=torch.int16, device=dev)
        self.val_pool = torch.zeros(max(val_len, 1), dtype=torch.float64, device=dev)
        self.partials = torch.zeros(nslots * self.width, dtype=torch.float64, device=dev)
        self.out = torch.zeros(ninst * self.width, dtype=torch.float64, device=dev)
        self.sync = torch.zeros(4, dtype=torch.int32, device=dev)
        self.angles = torch.zeros(ninst * 2 * p, dtype=torch.float64, device=dev)
        # cut tables on the GPU
        s = bk.stream_handle()
        for (n, edges, w, weighted, off) in cut_jobs:
            e = torch.from_numpy(edges.copy() if len(edges) else np.zeros((1, 2), np.int32)).to(dev)
            if weighted:
                wt = torch.from_numpy(w).to(dev)
                _native.call('qfx_cut_table', s, self._vp(self.val_pool, off * 8), n,
                             self._vp(e), self._vp(wt), len(edges))
            else:
                _native.call('qfx_cut_counts', s, self._vp(self.cnt_pool, off * 2), n,
                             self._vp(e), len(edges))
        for i, (n, edges, w, weighted, off) in enumerate(cut_jobs):
            tab = self.val_pool[off:off + 2 ** n] if weighted else \
                self.cnt_pool[off:off + 2 ** n]
            self.maxcut[i] = float(tab.max().item()) if weighted else \
                float((tab.to(torch.int32) & 0xFFFF).max().item())
        self._args = QaoaArgs()
        a = self._args
        a.insts, a.passes, a.items = self._insts.data_ptr(), self._passes.data_ptr(), \
            self._items.data_ptr()
        a.kstart, a.angles = self._kstart.data_ptr(), self.angles.data_ptr()
        a.psi_pool, a.lam_pool = self.psi_pool.data_ptr(), self.lam_pool.data_ptr()
        a.cnt_pool, a.val_pool = self.cnt_pool.data_ptr(), self.val_pool.data_ptr()
        a.partials, a.out, a.sync = self.partials.data_ptr(), self.out.data_ptr(), \
            self.sync.data_ptr()
        a.ninst, a.nk, a.width, a.mode, a.block = ninst, self.nk, self.width, mode, block
        a.max_items_per_k = max(kstart[k + 1] - kstart[k] for k in range(self.nk))
        a.tile_bits = T
        self._graph = None

    def algorithmic_bytes(self):
        """(state bytes, cut-table bytes) one launch moves, summed over the
        instances (program_bytes)."""
        st = ct = 0
        for g, n in zip(self.graphs, self.n):
            weighted = graph_arrays(g)[3]
            a, b = program_bytes(n, self.p, self.mode, self.T, self.block,
                                 self._mirror and use_mirror(n, self.T), weighted, self._fuse)
            st += a
            ct += b
        return st, ct

    @staticmethod
    def _vp(t, byte_off=0):
        return ctypes.c_void_p(t.data_ptr() + byte_off)

    @property
    def ninst(self):
        return len(self.graphs)

    def set_angles(self, beta, gamma):
        """beta, gamma: [ninst, p] arrays or tensors (copied to the device)."""
        b = torch.as_tensor(beta, dtype=torch.float64).reshape(self.ninst, self.p)
        g = torch.as_tensor(gamma, dtype=torch.float64).reshape(self.ninst, self.p)
        self.angles.copy_(torch.cat([b, g], dim=1).reshape(-1).to(self.angles.device))

    def launch(self):
        """Enqueue one evaluation on torch's current stream; results in
        ``self.out`` ([ninst, width]: <C>, then d/dbeta, d/dgamma)."""
        _native.call('qfx_qaoa_run', bk.stream_handle(), ctypes.byref(self._args))

    def capture(self):
        """Record launch() into a HIP graph (replayed by ``replay``)."""
        self.launch()                               # sets kernel attributes
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch()
        self._graph = g
        return g

    def replay(self):
        if self._graph is None:
            self.capture()
        self._graph.replay()

    def check(self):
        """Synchronise and raise if a grid barrier timed out in any launch
        since the batch was built (the kernel's error word is sticky: a
        later launch does not clear it)."""
        failed = ctypes.c_int(0)
        torch.cuda.synchronize()
        _native.call('qfx_qaoa_status', ctypes.c_void_p(self.sync.data_ptr()), ctypes.byref(failed))
        raise_if_failed(failed.value)

    def pass_times(self, max_grid=4096):
        """One launch with the kernel's wall-clock stamps on.  Per global pass
        k: its items, its span (first workgroup start to last arrival at the
        barrier after it) and the mean and largest time a workgroup spent on
        its items, in microseconds (device wall clock, qfx_init's rate).
        The gap between span and mean busy time is barrier wait and load
        imbalance."""
        dev = bk.DEVICE
        khz = bk.device_props()[7]
        if khz <= 0:
            raise RuntimeError('no device wall-clock rate')
        buf = torch.zeros(2 * self.nk * max_grid, dtype=torch.int64, device=dev)
        a = self._args
        a.timing, a.timing_len = buf.data_ptr(), buf.numel()
        try:
            self.launch()
            torch.cuda.synchronize()
        finally:
            a.timing, a.timing_len = None, 0
        self.check()
        flat = buf.cpu().numpy().astype(np.int64)
        # entries [k][workgroup][start, end] for the launch's grid
        grid = int(np.count_nonzero(flat)) // (2 * self.nk)
        st = flat[:2 * self.nk * grid].reshape(self.nk, grid, 2)
        us = 1e3 / khz
        ks = self._kstart.cpu().numpy()
        out = []
        for k in range(self.nk):
            s0, s1 = st[k, :, 0], st[k, :, 1]
            busy = (s1 - s0) * us
            out.append({'pass': k, 'items': int(ks[k + 1] - ks[k]),
                        'span_us': round(float(s1.max() - s0.min()) * us, 2),
                        'busy_mean_us': round(float(np.mean(busy)), 2),
                        'busy_max_us': round(float(np.max(busy)), 2)})
        total = float(st[:, :, 1].max() - st[0, :, 0].min()) * us
        return {'grid': grid, 'total_us': round(total, 2), 'passes': out}

    def evaluate(self, beta, gamma, graph=False):
        """(values [ninst], grads [ninst, 2p] or None) as device tensors.
        The launch is asynchronous: ``check()`` (which synchronises) raises
        if any launch since the batch was built hit a barrier timeout, so
        one call after a run of evaluations covers all of them."""
        self.set_angles(beta, gamma)
        if graph:
            self.replay()
        else:
            self.launch()
        out = self.out.view(self.ninst, self.width)
        if self.mode == MODE_FORWARD:
            return out[:, 0], None
        return out[:, 0], out[:, 1:]


def raise_if_failed(flag):
    """The engine's status word: nonzero means a grid barrier gave up, and
    the values and gradients of that launch are partial."""
    if flag:
        raise RuntimeError('qaoa engine: grid barrier timed out; results are partial')


class QAOAExpectation(torch.autograd.Function):
    """<C>(beta, gamma) for every instance of a QAOABatch; the backward pass
    uses the engine's adjoint gradients (computed in the same launch)."""

    @staticmethod
    def forward(ctx, beta, gamma, batch):
        val, grad = batch.evaluate(beta.detach(), gamma.detach())
        batch.check()
        ctx.p = batch.p
        ctx.save_for_backward(grad.clone() if grad is not None else val.new_zeros(0))
        return val.clone().to(beta.device)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        if grad.numel() == 0:
            raise RuntimeError('QAOABatch built in MODE_FORWARD has no gradients')
        g = grad.to(gout.device) * gout[:, None]
        return g[:, :ctx.p], g[:, ctx.p:], None


def qaoa_expectation(beta, gamma, batch):
    """Differentiable batched QAOA expectation (see QAOAExpectation)."""
    return QAOAExpectation.apply(beta, gamma, batch)


def optimize(batch, beta0, gamma0, learning_rate=0.01, max_steps=10000, block=10,
             min_gain=1e-4, use_graph=True, optimizer='gd'):
    """Gradient ascent on <C> for all instances at once, with the reference's
    stopping rule per instance (examples/qaoa_maxcut.py:63-82): 10-step
    blocks, an instance stops when its ratio <C>/maxcut gains less than
    1e-4 over a block.  optimizer 'gd' is the reference's plain gradient
    descent on -<C>; 'adam' uses Adam steps.  Returns (ratios, beta, gamma,
    steps per instance)."""
    if batch.mode != MODE_BACKWARD:
        raise ValueError('optimize needs a MODE_BACKWARD batch')
    dev = batch.angles.device
    beta = torch.as_tensor(beta0, dtype=torch.float64, device=dev).reshape(batch.ninst, batch.p).clone()
    gamma = torch.as_tensor(gamma0, dtype=torch.float64, device=dev).reshape(batch.ninst, batch.p).clone()
    maxcut = torch.as_tensor(batch.maxcut, dtype=torch.float64, device=dev)
    active = torch.ones(batch.ninst, dtype=torch.bool, device=dev)
    steps = torch.zeros(batch.ninst, dtype=torch.int64, device=dev)
    last = torch.full((batch.ninst,), -1.0, dtype=torch.float64, device=dev)
    ratio = torch.zeros(batch.ninst, dtype=torch.float64, device=dev)
    m = torch.zeros(batch.ninst, 2 * batch.p, dtype=torch.float64, device=dev)
    v = torch.zeros_like(m)
    t = 0
    for step in range(0, max_steps, block):
        for _ in range(block):
            val, grad = batch.evaluate(beta, gamma, graph=use_graph)
            if optimizer == 'adam':
                t += 1
                m.mul_(0.9).add_(0.1 * grad)
                v.mul_(0.999).add_(0.001 * grad * grad)
                upd = learning_rate * (m / (1 - 0.9 ** t)) / (torch.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
            else:
                upd = learning_rate * grad
            upd = upd * active[:, None]
            beta = beta + upd[:, :batch.p]
            gamma = gamma + upd[:, batch.p:]
            steps += active.to(torch.int64)
        val, _ = batch.evaluate(beta, gamma, graph=use_graph)
        batch.check()           # every launch of the block, before its values are used
        ratio = torch.where(active, val / maxcut, ratio)
        done = (ratio - last) < min_gain
        active = active & ~done
        last = torch.where(active, ratio, last)
        if not bool(active.any()):
            break
    return ratio, beta, gamma, steps

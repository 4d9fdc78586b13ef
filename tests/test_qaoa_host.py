"""Host checks of the batched QAOA engine's programs (no GPU).

``emulate`` runs a program's stages, in pass order, on a full numpy state
with the stage semantics the kernel implements.  Matching the oracle's
gate-by-gate <C> and its exact parameter-shift gradient pins the pass
scheduling (every RX bit once per layer, cost layers in order) and the
adjoint formulas
  d<C>/dbeta_l  = Im sum_j <lambda|X_j|psi>,
  d<C>/dgamma_l = 2 Im <lambda|C|psi> - W Im <lambda|psi>.
"""
import networkx as nx
import numpy as np
import pytest

from oracle import qf_oracle as O
from quantumflow_amd import qaoa_engine as E


def emulate(n, edges, weights, beta, gamma, mode, T, fuse=True, block=512, mirror=False):
    cuts = O.graph_cuts(n, edges, weights).reshape(-1)
    W = float(len(edges) if weights is None else np.sum(weights))
    p = len(beta)
    prog = E.build_program(n, p, mode, T, fuse, block, mirror=mirror)
    idx = np.arange(2 ** n)
    psi = lam = None
    value = 0.0
    gb = np.zeros(p)
    gg = np.zeros(p)

    def rx(v, bits, theta):
        c, s = np.cos(theta / 2), np.sin(theta / 2)
        for b in bits:
            lo = v[(idx >> b) & 1 == 0]
            hi = v[(idx >> b) & 1 == 1]
            v = v.copy()
            v[(idx >> b) & 1 == 0] = c * lo - 1j * s * hi
            v[(idx >> b) & 1 == 1] = c * hi - 1j * s * lo
        return v

    seen = {}
    for tset, stages, flags in prog:
        for typ, layer, bits in stages:
            if bits is not None:
                assert set(bits) <= set(tset)
            if typ == E.ST_INIT:
                psi = np.full(2 ** n, 2.0 ** (-n / 2), dtype=complex)
            elif typ == E.ST_COST:
                psi = psi * np.exp(1j * gamma[layer] * (W / 2 - cuts))
            elif typ == E.ST_COST_RX:
                # phases first, then the mixer chunk(s) of the same layer
                psi = psi * np.exp(1j * gamma[layer] * (W / 2 - cuts))
                for b in bits:
                    seen[(layer, b)] = seen.get((layer, b), 0) + 1
                psi = rx(psi, sorted(bits), beta[layer])
            elif typ in (E.ST_RX, E.ST_RX_LOAD, E.ST_RX_STORE):
                if typ == E.ST_RX_LOAD:
                    assert stages[0][0] == typ and flags & E.PF_LOAD_PSI
                if typ == E.ST_RX_STORE:
                    assert stages[-1][0] == typ and flags & E.PF_STORE_PSI
                for b in bits:
                    seen[(layer, b)] = seen.get((layer, b), 0) + 1
                psi = rx(psi, sorted(bits), beta[layer])
            elif typ == E.ST_EXPECT:
                value += np.sum(np.abs(psi) ** 2 * cuts)
            elif typ == E.ST_MAKE_LAMBDA:
                lam = cuts * psi
            elif typ == E.ST_RX_ADJ:
                for b in bits:
                    partner = psi[idx ^ (1 << b)]
                    gb[layer] += np.sum(np.conj(lam) * partner).imag
                psi = rx(psi, sorted(bits), -beta[layer])
                lam = rx(lam, sorted(bits), -beta[layer])
            elif typ == E.ST_COST_ADJ:
                lp = np.conj(lam) * psi
                gg[layer] += np.sum(2 * cuts * lp.imag - W * lp.imag)
                ph = np.exp(-1j * gamma[layer] * (W / 2 - cuts))
                psi, lam = psi * ph, lam * ph
    assert all(seen.get((l, b), 0) == 1 for l in range(p) for b in range(n))
    return value, gb, gg


@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("n,T,p", [(5, 12, 2), (7, 4, 2), (8, 5, 3), (9, 6, 1)])
def test_program_emulation_matches_oracle(n, T, p, fuse):
    rng = np.random.default_rng(n * 10 + T)
    g = nx.gnp_random_graph(n, 0.5, seed=int(rng.integers(1000)))
    edges = list(g.edges())
    beta = rng.uniform(0, np.pi / 2, size=p)
    gamma = rng.uniform(0, np.pi, size=p)
    ref = O.qaoa_expectation(n, edges, beta, gamma)
    v, gb, gg = emulate(n, edges, None, beta, gamma, E.MODE_BACKWARD, T, fuse)
    assert abs(v - ref) < 1e-12
    sb, sg = O.qaoa_grad_shift(n, edges, beta, gamma)
    assert np.max(np.abs(gb - sb)) < 1e-12
    assert np.max(np.abs(gg - sg)) < 1e-12
    v2, _, _ = emulate(n, edges, None, beta, gamma, E.MODE_FORWARD, T + 1, fuse)
    assert abs(v2 - ref) < 1e-12
    if E.use_mirror(n, T):
        # the half-state schedules (every set holding bit n-1) are complete too
        vm, gbm, ggm = emulate(n, edges, None, beta, gamma, E.MODE_BACKWARD, T, fuse, mirror=True)
        assert abs(vm - ref) < 1e-12
        assert np.max(np.abs(gbm - sb)) < 1e-12 and np.max(np.abs(ggm - sg)) < 1e-12
    # small workgroups put register-held RX bits inside these small tiles
    for block in (2, 4, 8):
        v3, _, _ = emulate(n, edges, None, beta, gamma, E.MODE_FORWARD, T, fuse, block)
        assert abs(v3 - ref) < 1e-12


def test_weighted_program_gradient():
    rng = np.random.default_rng(3)
    n, p = 6, 2
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (0, 5), (1, 4)]
    w = list(rng.uniform(0.2, 2.0, size=len(edges)))
    beta = rng.uniform(0, 1, size=p)
    gamma = rng.uniform(0, 1, size=p)
    v, gb, gg = emulate(n, edges, w, beta, gamma, E.MODE_BACKWARD, 4)
    assert abs(v - O.qaoa_expectation(n, edges, beta, gamma, w)) < 1e-12
    sb, sg = O.qaoa_grad_shift(n, edges, beta, gamma, w)
    assert np.max(np.abs(gb - sb)) < 1e-12 and np.max(np.abs(gg - sg)) < 1e-12


@pytest.mark.parametrize("n", [14, 20, 22, 24])
def test_pass_counts(n):
    fwd = E.build_program(n, 4, E.MODE_FORWARD, 13)
    bwd = E.build_program(n, 4, E.MODE_BACKWARD, 12)
    assert len(fwd) <= 2 * 4 + 1
    assert len(bwd) <= 4 * 4 + 1
    for prog in (fwd, bwd):
        assert prog[0][2] & E.PF_LOAD_PSI == 0          # first pass builds |+>
        assert prog[-1][2] & E.PF_STORE_PSI == 0        # last pass stores nothing
        for tset, stages, flags in prog:
            assert len(tset) <= 13 and list(tset) == sorted(tset)


@pytest.mark.parametrize("n,T", [(9, 12), (20, 13), (24, 12)])
def test_fused_stages(n, T):
    """In forward programs every cost stage merges into the mixer stage of
    its layer that follows it (chained passes always hold one), with a
    non-empty bit set; backward programs stay unfused."""
    p = 3
    for mode in (E.MODE_FORWARD, E.MODE_BACKWARD):
        plain = E.build_program(n, p, mode, T, fuse=False)
        fused = E.build_program(n, p, mode, T)
        assert [x[0] for x in plain] == [x[0] for x in fused]
        assert [x[2] for x in plain] == [x[2] for x in fused]
        ncost = sum(st[0] == E.ST_COST for _, sts, _ in plain for st in sts)
        nf = sum(st[0] == E.ST_COST_RX for _, sts, _ in fused for st in sts)
        assert ncost == p
        if mode == E.MODE_FORWARD:
            assert nf == p
        else:
            assert nf == 0 and plain == fused
        for _, sts, _ in fused:
            for typ, layer, bits in sts:
                if typ == E.ST_COST_RX:
                    assert bits


def test_tile_sets_cover():
    for n in range(1, 27):
        for T in (12, 13):
            sets = E.tile_sets(n, T)
            assert set().union(*map(set, sets)) == set(range(n))
            assert all(len(s) == min(T, n) for s in sets)


class _FakeBatch:
    """A CPU stand-in for QAOABatch: <C> = -|x|^2 per instance and a status
    word that a chosen launch sets and no later launch clears (the kernel's
    sticky error flag)."""
    mode = E.MODE_BACKWARD

    def __init__(self, ninst, p, fail_at):
        import torch
        self.ninst, self.p = ninst, p
        self.angles = torch.zeros(ninst * 2 * p, dtype=torch.float64)
        self.maxcut = np.ones(ninst)
        self.fail_at, self.launches, self.flag, self.checks = fail_at, 0, 0, 0

    def evaluate(self, beta, gamma, graph=False):
        import torch
        self.launches += 1
        if self.launches == self.fail_at:
            self.flag = 1
        x = torch.cat([torch.as_tensor(beta), torch.as_tensor(gamma)], dim=1)
        return 1.0 - (x * x).sum(1), -2 * x

    def check(self):
        self.checks += 1
        E.raise_if_failed(self.flag)


def test_status_flag_raises():
    E.raise_if_failed(0)
    with pytest.raises(RuntimeError, match="grid barrier"):
        E.raise_if_failed(1)


def test_optimize_sees_a_timeout_from_inside_a_block():
    # the failing launch is the 3rd of the first 10-step block; later
    # launches do not clear the flag, and the block-end check raises
    b = _FakeBatch(3, 2, fail_at=3)
    with pytest.raises(RuntimeError, match="grid barrier"):
        E.optimize(b, np.full((3, 2), 0.5), np.full((3, 2), 0.5), use_graph=False)
    assert b.checks == 1 and b.launches == 11
    ok = _FakeBatch(3, 2, fail_at=-1)
    ratio, beta, gamma, steps = E.optimize(ok, np.full((3, 2), 0.5), np.full((3, 2), 0.5),
                                           use_graph=False, learning_rate=0.1)
    assert ok.checks >= 1 and np.all(ratio.numpy() > 0.9)


def test_autograd_forward_checks_status():
    import torch
    b = _FakeBatch(2, 1, fail_at=1)
    beta = torch.full((2, 1), 0.1, dtype=torch.float64, requires_grad=True)
    gamma = torch.full((2, 1), 0.2, dtype=torch.float64, requires_grad=True)
    with pytest.raises(RuntimeError, match="grid barrier"):
        E.qaoa_expectation(beta, gamma, b)
    ok = _FakeBatch(2, 1, fail_at=-1)
    v = E.qaoa_expectation(beta, gamma, ok)
    v.sum().backward()
    assert ok.checks == 1
    assert torch.allclose(beta.grad, -2 * beta.detach())


def test_program_bytes_closed_form():
    """program_bytes: a forward program of p + 1 passes loads and stores the
    state between passes, 2 p state transfers of 16 B per amplitude; a
    mirror instance with the same pass count moves half; a one-pass instance
    only reads its cut counts (2 B per amplitude per cost-reading stage)."""
    from quantumflow_amd import qaoa_engine as E
    n, p, T = 22, 4, 13
    prog = E.build_program(n, p, E.MODE_FORWARD, T)
    assert len(prog) == p + 1
    st, ct = E.program_bytes(n, p, E.MODE_FORWARD, T)
    assert st == 16 * 2 ** n * 2 * p
    reads = sum(s[0] in (E.ST_COST, E.ST_COST_RX, E.ST_EXPECT) for _, stages, _ in prog
                for s in stages)
    assert reads == p + 1 and ct == 2 * 2 ** n * (p + 1)
    mprog = E.build_program(n, p, E.MODE_FORWARD, T, mirror=True)
    assert len(mprog) == len(prog)
    assert all(tset[-1] == n - 1 and len(tset) == T for tset, _, _ in mprog)
    mst, mct = E.program_bytes(n, p, E.MODE_FORWARD, T, mirror=True)
    assert (mst, mct) == (st // 2, ct // 2)
    st1, ct1 = E.program_bytes(10, p, E.MODE_FORWARD, T)
    assert st1 == 0 and ct1 == 2 * 2 ** 10 * (p + 1)
    assert E.use_mirror(T + 2, T) and not E.use_mirror(T + 1, T)


def _emulate_mirror_tiles(n, edges, beta, gamma, T):
    """A forward mirror program run tile by tile as the kernel runs it: the
    state kept is the half with index bit n-1 clear; each item of a pass
    gathers its tile (the pass's tile bits over the item's non-tile bits,
    whose highest is clear), an element with bit n-1 set from the
    complementary index; stages act on the tile; the tile is scattered back
    the same way.  <C> sums every item twice."""
    cuts = O.graph_cuts(n, edges).reshape(-1)
    W = float(len(edges))
    p = len(beta)
    prog = E.build_program(n, p, E.MODE_FORWARD, T, fuse=False, mirror=True)
    full = (1 << n) - 1
    kept = np.zeros(2 ** (n - 1), dtype=complex)

    def stored(z):
        return z ^ full if (z >> (n - 1)) & 1 else z

    value = 0.0
    for tset, stages, flags in prog:
        assert tset[-1] == n - 1 and len(tset) == T
        others = [b for b in range(n) if b not in tset]
        for tile in range(2 ** (len(others) - 1)):
            base = sum(((tile >> j) & 1) << b for j, b in enumerate(others))
            zs = [base | sum(((e >> j) & 1) << b for j, b in enumerate(tset)) for e in range(2 ** T)]
            at = np.array([stored(z) for z in zs])
            v = kept[at].copy()
            c = cuts[at]
            for typ, layer, bits in stages:
                if typ == E.ST_INIT:
                    v[:] = 2.0 ** (-n / 2)
                elif typ == E.ST_COST:
                    v = v * np.exp(1j * gamma[layer] * (W / 2 - c))
                elif typ == E.ST_RX:
                    cs, sn = np.cos(beta[layer] / 2), np.sin(beta[layer] / 2)
                    for b in sorted(bits):
                        j = tset.index(b)
                        e = np.arange(2 ** T)
                        lo, hi = e[(e >> j) & 1 == 0], e[(e >> j) & 1 == 1]
                        x, y = v[lo].copy(), v[hi].copy()
                        v[lo], v[hi] = cs * x - 1j * sn * y, cs * y - 1j * sn * x
                elif typ == E.ST_EXPECT:
                    value += 2 * np.sum(np.abs(v) ** 2 * c)
            kept[at] = v
    return value


@pytest.mark.parametrize("n,T,p", [(8, 6, 2), (9, 6, 2), (10, 7, 3)])
def test_mirror_program_on_tiles_matches_oracle(n, T, p):
    """The half-state (mirror) layout, run tile by tile on the host as the
    kernel runs it, gives the oracle's <C>; its stage schedule (full-state
    emulation) covers every RX bit once per layer."""
    rng = np.random.default_rng(n * 7 + T)
    g = nx.gnp_random_graph(n, 0.5, seed=int(rng.integers(1000)))
    edges = list(g.edges())
    beta = rng.uniform(0, np.pi / 2, size=p)
    gamma = rng.uniform(0, np.pi, size=p)
    ref = O.qaoa_expectation(n, edges, beta, gamma)
    assert E.use_mirror(n, T)
    assert abs(_emulate_mirror_tiles(n, edges, beta, gamma, T) - ref) < 1e-12

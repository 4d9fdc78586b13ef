"""Learned QAOA angle proposer: a small MLP trained and run by the
package's own fp32 HIP kernels (csrc/mlp.hip).

Inputs (19 features of a graph): n, |E|, the degree histogram in 16 bins
(degrees >= 15 share the last bin) and the triangle count.  Output: the 2p
angles (beta_0..beta_{p-1}, gamma_0..gamma_{p-1}) used to start the QAOA
optimizer.  Architecture 19 -> 256 -> 256 -> 256 -> 2p with ReLU; the output
bias starts at the linear-ramp angles, so an untrained proposer returns the
fallback initialisation.

The reference has no learned component; the proposer warm-starts the
reference's optimizer loop (examples/qaoa_maxcut.py:37-87), whose plain
start (here the linear ramp) stays the fallback and the simulated
expectation the judge of every proposal (``evaluate``).
"""

import ctypes

import networkx as nx
import numpy as np
import torch

from . import _native
from . import backend as bk

FEATURES = 19
HIDDEN = (256, 256, 256)


def graph_features(graph, normalise=True):
    """19 features of a graph (see module docstring)."""
    n = graph.number_of_nodes()
    m = graph.number_of_edges()
    hist = np.zeros(16)
    for _, d in graph.degree():
        hist[min(d, 15)] += 1
    tri = sum(nx.triangles(graph).values()) // 3
    f = np.concatenate([[n, m], hist, [tri]]).astype(np.float64)
    if normalise:
        f = f.copy()
        f[0] /= 20.0
        f[1] /= 100.0
        f[2:18] /= max(n, 1)
        f[18] /= 100.0
    return f


def linear_ramp(p, delta=0.75):
    """beta_i = (1 - (i+1/2)/p) delta, gamma_i = (i+1/2)/p delta."""
    f = (np.arange(p) + 0.5) / p
    return (1.0 - f) * delta, f * delta


def _vp(t, off_elems=0):
    return ctypes.c_void_p(t.data_ptr() + off_elems * t.element_size())


class AngleProposer:
    """MLP 19 -> hidden -> 2p (fp32) with parameters in one device buffer."""

    def __init__(self, p=4, hidden=HIDDEN, seed=0, ramp_delta=0.75):
        self.p = p
        self.sizes = [FEATURES] + list(hidden) + [2 * p]
        rng = np.random.default_rng(seed)
        chunks = []
        self.layout = []                         # (W offset, b offset, in, out)
        off = 0
        for li, (fi, fo) in enumerate(zip(self.sizes[:-1], self.sizes[1:])):
            last = li == len(self.sizes) - 2
            lim = np.sqrt(6.0 / fi) * (0.1 if last else 1.0)
            W = rng.uniform(-lim, lim, size=(fo, fi))
            b = np.zeros(fo)
            if last:
                rb, rg = linear_ramp(p, ramp_delta)
                b = np.concatenate([rb, rg])
            self.layout.append((off, off + fo * fi, fi, fo))
            chunks += [W.reshape(-1), b]
            off += fo * fi + fo
        self.nparams = off
        dev = bk.DEVICE
        bk.stream_handle()
        self.params = torch.from_numpy(np.concatenate(chunks).astype(np.float32)).to(dev)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0

    # ---- kernels --------------------------------------------------------
    def _gemm(self, C, A, B, M, N, K, sam, sak, sbk, sbn, bias=None, relu=False,
              boff=0, bias_off=0, aoff=0, coff=0):
        _native.call('qfx_mlp_gemm', bk.stream_handle(), _vp(C, coff), _vp(A, aoff),
                     _vp(B, boff), M, N, K, sam, sak, sbk, sbn,
                     None if bias is None else _vp(bias, bias_off), int(relu))

    def forward(self, x, keep=False):
        """x: [B, 19] float32 device tensor -> [B, 2p] (and the activations)."""
        x = x.contiguous()
        B = x.shape[0]
        acts = [x]
        h = x
        for li, (wo, bo, fi, fo) in enumerate(self.layout):
            y = torch.empty(B, fo, dtype=torch.float32, device=x.device)
            last = li == len(self.layout) - 1
            self._gemm(y, h, self.params, B, fo, fi, fi, 1, 1, fi, bias=self.params,
                       relu=not last, boff=wo, bias_off=bo)
            acts.append(y)
            h = y
        return (h, acts) if keep else h

    def backward(self, acts, dy):
        """Parameter gradients into self.grads for d loss / d output = dy."""
        B = dy.shape[0]
        s = bk.stream_handle()
        g = dy.contiguous()
        for li in reversed(range(len(self.layout))):
            wo, bo, fi, fo = self.layout[li]
            x = acts[li]
            if li < len(self.layout) - 1:
                dz = torch.empty_like(g)
                _native.call('qfx_mlp_relu_bwd', s, _vp(dz), _vp(g), _vp(acts[li + 1]), g.numel())
                g = dz
            # dW = dz^T x ; db = column sums of dz
            self._gemm(self.grads, g, x, fo, fi, B, 1, fo, fi, 1, coff=wo)
            _native.call('qfx_mlp_colsum', s, _vp(self.grads, bo), _vp(g), B, fo)
            if li > 0:
                dx = torch.empty(B, fi, dtype=torch.float32, device=g.device)
                self._gemm(dx, g, self.params, B, fi, fo, fo, 1, fi, 1, boff=wo)
                g = dx

    def loss_and_grad(self, x, t):
        """MSE loss (0-dim device tensor) and gradients in self.grads."""
        y, acts = self.forward(x, keep=True)
        dy = torch.empty_like(y)
        loss = torch.empty((), dtype=torch.float32, device=y.device)
        ws = torch.empty(max(1, _native.load().qfx_mlp_mse_workspace(y.numel()) // 4),
                         dtype=torch.float32, device=y.device)
        _native.call('qfx_mlp_mse', bk.stream_handle(), _vp(dy), _vp(loss), _vp(y),
                     _vp(t.contiguous()), y.numel(), _vp(ws))
        self.backward(acts, dy)
        return loss

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self.step += 1
        _native.call('qfx_mlp_adam', bk.stream_handle(), _vp(self.params), _vp(self.grads),
                     _vp(self.m), _vp(self.v), self.nparams, lr, beta1, beta2, eps, self.step,
                     weight_decay)

    def fit(self, X, T, epochs=200, batch=512, lr=1e-3, seed=0, weight_decay=0.0,
            verbose=False):
        """Mini-batch Adam on (features, target angles); returns per-epoch
        mean losses (host floats)."""
        X = torch.as_tensor(X, dtype=torch.float32).to(self.params.device)
        T = torch.as_tensor(T, dtype=torch.float32).to(self.params.device)
        N = X.shape[0]
        gen = torch.Generator().manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(N, generator=gen).to(X.device)
            tot = torch.zeros((), dtype=torch.float32, device=X.device)
            nb = 0
            for i in range(0, N, batch):
                idx = perm[i:i + batch]
                tot += self.loss_and_grad(X[idx], T[idx])
                self.adam_step(lr, weight_decay=weight_decay)
                nb += 1
            hist.append(float((tot / nb).item()))
            if verbose:
                print('epoch', ep, 'loss', hist[-1], flush=True)
        return hist

    def propose(self, graphs):
        """(beta [N, p], gamma [N, p]) numpy arrays for a list of graphs."""
        X = np.stack([graph_features(g) for g in graphs]).astype(np.float32)
        y = self.forward(torch.from_numpy(X).to(self.params.device))
        y = y.double().cpu().numpy()
        return y[:, :self.p], y[:, self.p:]

    def state_dict(self):
        return {'p': self.p, 'sizes': np.array(self.sizes), 'params': self.params.cpu().numpy()}

    def save(self, path):
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path):
        d = np.load(path)
        sizes = [int(x) for x in d['sizes']]
        obj = cls(p=int(d['p']), hidden=tuple(sizes[1:-1]))
        obj.params.copy_(torch.from_numpy(d['params']).to(obj.params.device))
        return obj


# ---- data and evaluation ----------------------------------------------------

def random_graphs(count, n_lo=8, n_hi=20, pe_lo=0.3, pe_hi=0.6, seed=0):
    """ER graphs, n ~ U{n_lo..n_hi}, edge probability ~ U[pe_lo, pe_hi];
    graphs without edges are redrawn."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        n = int(rng.integers(n_lo, n_hi + 1))
        g = nx.gnp_random_graph(n, float(rng.uniform(pe_lo, pe_hi)), seed=int(rng.integers(1 << 30)))
        if g.number_of_edges() > 0:
            out.append(g)
    return out


def optimise_graphs(graphs, p, beta0, gamma0, steps, lr=0.02, chunk=1024, optimizer='adam',
                    trace_every=0, progress=False):
    """Run the batched optimizer on graphs in chunks; returns final
    (beta, gamma, ratio) and, with trace_every > 0, the ratio after every
    trace_every steps ([N, steps // trace_every + 1], column 0 = start)."""
    from . import qaoa_engine as E
    N = len(graphs)
    beta = np.zeros((N, p))
    gamma = np.zeros((N, p))
    ratio = np.zeros(N)
    trace = None
    if trace_every:
        trace = np.zeros((N, steps // trace_every + 1))
    for c0 in range(0, N, chunk):
        sub = graphs[c0:c0 + chunk]
        batch = E.QAOABatch(sub, p, mode=E.MODE_BACKWARD)
        mc = torch.as_tensor(batch.maxcut, dtype=torch.float64, device=batch.angles.device)
        b = torch.as_tensor(beta0[c0:c0 + chunk], dtype=torch.float64, device=mc.device).clone()
        g = torch.as_tensor(gamma0[c0:c0 + chunk], dtype=torch.float64, device=mc.device).clone()
        m = torch.zeros(len(sub), 2 * p, dtype=torch.float64, device=mc.device)
        v = torch.zeros_like(m)
        batch.set_angles(b, g)
        batch.capture()
        for t in range(1, steps + 1):
            batch.set_angles(b, g)
            batch.replay()
            out = batch.out.view(len(sub), -1)
            if trace is not None and (t - 1) % trace_every == 0:
                trace[c0:c0 + len(sub), (t - 1) // trace_every] = (out[:, 0] / mc).cpu().numpy()
            grad = out[:, 1:]
            if optimizer == 'adam':
                m.mul_(0.9).add_(0.1 * grad)
                v.mul_(0.999).add_(0.001 * grad * grad)
                upd = lr * (m / (1 - 0.9 ** t)) / (torch.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
            else:
                upd = lr * grad
            b = b + upd[:, :p]
            g = g + upd[:, p:]
        batch.set_angles(b, g)
        batch.replay()
        out = batch.out.view(len(sub), -1)
        r = (out[:, 0] / mc).cpu().numpy()
        if trace is not None and steps % trace_every == 0:
            trace[c0:c0 + len(sub), -1] = r
        batch.check()
        beta[c0:c0 + len(sub)] = b.cpu().numpy()
        gamma[c0:c0 + len(sub)] = g.cpu().numpy()
        ratio[c0:c0 + len(sub)] = r
        if progress:
            print('optimised graphs %d..%d of %d, mean ratio %.4f'
                  % (c0, c0 + len(sub), N, float(np.mean(r))), flush=True)
    return beta, gamma, ratio, trace


def expectation_ratios(graphs, p, beta, gamma, chunk=2048):
    """<C>/maxcut at the given angles (maxcut = exact brute force over all
    assignments, the cut table's maximum)."""
    from . import qaoa_engine as E
    out = np.zeros(len(graphs))
    for c0 in range(0, len(graphs), chunk):
        sub = graphs[c0:c0 + chunk]
        batch = E.QAOABatch(sub, p, mode=E.MODE_FORWARD)
        val, _ = batch.evaluate(beta[c0:c0 + chunk], gamma[c0:c0 + chunk])
        out[c0:c0 + len(sub)] = val.cpu().numpy() / batch.maxcut
    return out


def run_pipeline(p=4, ntrain=20000, neval=1000, n_lo=8, n_hi=20, opt_steps=200, lr=0.02,
                 epochs=100, batch=512, seed=0, hidden=HIDDEN, verbose=True):
    """Learned warm start end to end: optimise ntrain graphs from the linear
    ramp with the batched engine, train the proposer on (features,
    optimised angles), then on neval fresh graphs compare proposer vs
    linear-ramp starts: starting ratio <C>/maxcut, final ratio after
    opt_steps optimizer steps, and steps to reach 0.95, 0.99 and 0.999 x the linear-ramp
    run's final ratio (maxcut = exact brute force)."""
    import time
    log = (lambda *a: print(*a, flush=True)) if verbose else (lambda *a: None)
    res = {'p': p, 'ntrain': ntrain, 'neval': neval, 'n_range': [n_lo, n_hi],
           'optimizer': 'adam lr=%g, %d steps' % (lr, opt_steps)}
    t0 = time.perf_counter()
    train = random_graphs(ntrain, n_lo, n_hi, seed=seed)
    rb, rg = linear_ramp(p)
    bopt, gopt, ratio, _ = optimise_graphs(train, p, np.tile(rb, (ntrain, 1)),
                                           np.tile(rg, (ntrain, 1)), opt_steps, lr=lr,
                                           progress=verbose)
    res['data_seconds'] = round(time.perf_counter() - t0, 2)
    res['train_mean_optimised_ratio'] = float(np.mean(ratio))
    log('dataset', res['data_seconds'], 's, mean ratio', res['train_mean_optimised_ratio'])
    X = np.stack([graph_features(g) for g in train]).astype(np.float32)
    T = np.concatenate([bopt, gopt], axis=1).astype(np.float32)
    t0 = time.perf_counter()
    prop = AngleProposer(p=p, hidden=hidden, seed=seed + 1)
    hist = prop.fit(X, T, epochs=epochs, batch=batch, lr=1e-3, seed=seed, verbose=verbose)
    res['train_seconds'] = round(time.perf_counter() - t0, 2)
    res['train_loss_first_last'] = [hist[0], hist[-1]]
    log('training', res['train_seconds'], 's, loss', hist[0], '->', hist[-1])
    test = random_graphs(neval, n_lo, n_hi, seed=seed + 1000)
    pb, pg = prop.propose(test)
    starts = {'proposer': (pb, pg),
              'linear_ramp': (np.tile(rb, (neval, 1)), np.tile(rg, (neval, 1)))}
    traces = {}
    for name, (b0, g0) in starts.items():
        _, _, fin, tr = optimise_graphs(test, p, b0, g0, opt_steps, lr=lr, trace_every=1,
                                        progress=verbose)
        traces[name] = tr
        res[name] = {'start_ratio_mean': float(np.mean(tr[:, 0])),
                     'final_ratio_mean': float(np.mean(fin))}
    for frac in (0.95, 0.99, 0.999):
        target = frac * traces['linear_ramp'][:, -1]
        for name, tr in traces.items():
            hit = tr >= target[:, None]
            steps = np.where(hit.any(axis=1), hit.argmax(axis=1), tr.shape[1])
            key = 'steps_to_%g_of_ramp_optimum' % frac
            res[name][key + '_mean'] = float(np.mean(steps))
            res[name][key + '_median'] = float(np.median(steps))
            res[name][key + '_reached_fraction'] = float(np.mean(hit.any(axis=1)))
    res['proposer_wins_start_ratio_fraction'] = float(np.mean(traces['proposer'][:, 0] >
                                                              traces['linear_ramp'][:, 0]))
    log(res)
    return res, prop


if __name__ == '__main__':
    import argparse
    import json
    ap = argparse.ArgumentParser(description='train and evaluate the QAOA angle proposer')
    ap.add_argument('--train', type=int, default=20000)
    ap.add_argument('--eval', type=int, default=1000)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--epochs', type=int, default=100)
    ap.add_argument('--out', default='')
    ap.add_argument('--model', default='')
    a = ap.parse_args()
    r, model = run_pipeline(ntrain=a.train, neval=a.eval, opt_steps=a.steps, epochs=a.epochs)
    if a.out:
        with open(a.out, 'w') as f:
            json.dump(r, f, indent=1)
    if a.model:
        model.save(a.model)

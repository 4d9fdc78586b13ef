"""float64 numpy restatement of the angle-proposer MLP (TEST ORACLE).

Test infrastructure only.  Dense layers y = relu(x W^T + b) (no ReLU on the
last layer), mean-squared-error loss, manual backward pass and Adam
(Kingma & Ba, ICLR 2015, Algorithm 1, with L2 weight decay added to the
gradient).  The reference has no learned component; this restates the
network the build's kernels implement.  Pinned in tests/test_oracle_mlp.py by
central finite differences of the loss and by Adam's closed-form first step.
"""

import numpy as np


def forward(params, x):
    """params: list of (W [out, in], b [out]); returns the activations of
    every layer (the input first)."""
    acts = [x]
    h = x
    for i, (W, b) in enumerate(params):
        z = h @ W.T + b
        h = np.maximum(z, 0.0) if i < len(params) - 1 else z
        acts.append(h)
    return acts


def mse(y, t):
    d = y - t
    return float(np.mean(d * d)), 2.0 * d / d.size


def backward(params, acts, dy):
    """Gradients [(dW, db)] of a scalar loss with d loss / d output = dy."""
    grads = [None] * len(params)
    g = dy
    for i in reversed(range(len(params))):
        W, b = params[i]
        x = acts[i]
        if i < len(params) - 1:
            g = g * (acts[i + 1] > 0)
        grads[i] = (g.T @ x, g.sum(axis=0))
        g = g @ W
    return grads


def loss_and_grads(params, x, t):
    acts = forward(params, x)
    loss, dy = mse(acts[-1], t)
    return loss, backward(params, acts, dy)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    g = g + wd * p
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v

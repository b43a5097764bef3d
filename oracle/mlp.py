"""Float64 numpy restatement of the PDT-surrogate MLP (forward, loss, backward).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The surrogate has no
pyAtmosphere counterpart; this oracle is pinned by central finite differences
in tests/test_mlp_oracle.py, and the device kernels are checked against it.

Layers: h_{l+1} = SiLU(h_l W_l^T + b_l) for hidden layers, linear output of
NB logits + 1 log-variance.  Loss per row: -sum_j t_j log p_j +
wc sum_j (P_j - T_j)^2 + wv (lv - tv)^2 with p = softmax(logits) and P, T the
cumulative sums of p and t, averaged over rows.
"""
import numpy as np


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def forward(params, x):
    """params: list of (W [out, in], b [out]); returns output and caches."""
    h = x
    cache = []
    for i, (W, b) in enumerate(params):
        z = h @ W.T + b
        cache.append((h, z))
        h = silu(z) if i < len(params) - 1 else z
    return h, cache


def loss_and_grad(params, x, target, tvar, wv, wc=0.0):
    out, cache = forward(params, x)
    M, NB = target.shape
    logits, lv = out[:, :NB], out[:, NB]
    mx = logits.max(axis=1, keepdims=True)
    lse = mx + np.log(np.exp(logits - mx).sum(axis=1, keepdims=True))
    logp = logits - lse
    p = np.exp(logp)
    dc = np.cumsum(p, axis=1) - np.cumsum(target, axis=1)
    loss = (-(target * logp).sum(axis=1) + wc * (dc ** 2).sum(axis=1) + wv * (lv - tvar) ** 2).mean()
    g = 2 * wc * np.cumsum(dc[:, ::-1], axis=1)[:, ::-1]      # d/dp_i of the CDF term
    dout = np.zeros_like(out)
    dout[:, :NB] = (p - target + p * (g - (p * g).sum(axis=1, keepdims=True))) / M
    dout[:, NB] = 2 * wv * (lv - tvar) / M
    grads = [None] * len(params)
    d = dout
    for i in reversed(range(len(params))):
        h, z = cache[i]
        if i < len(params) - 1:
            d = d * dsilu(z)
        W, b = params[i]
        grads[i] = (d.T @ h, d.sum(axis=0))
        d = d @ W
    return loss, grads

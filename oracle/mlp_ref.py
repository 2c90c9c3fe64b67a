"""CPU restatement of the schedule cost-model network (test infrastructure only).

Used by tests/ to check the library's HIP MLP kernels (forward, backward, Adam) and nothing
else; no product code imports it.  The network has no counterpart in the reference (which fixes
its MSM window, lib/cbits/curves/g1/proj/bn128_G1_proj.c:600); its definition is the plain
textbook one, computed here in float64:

    h_0 = x,  h_{l+1} = relu(W_l h_l + b_l) for hidden layers, y = W_{L-1} h_{L-1} + b_{L-1}
    loss = (1/B) sum_b |y_b - t_b|^2   (B rows; the mean squared error for one output)
    Adam (Kingma & Ba 2015, Algorithm 1) with bias-corrected moments.

Parameters are one flat vector laid out W_0 (out x in, row-major), b_0, W_1, b_1, ...
The analytic gradient here is pinned against central finite differences in the CPU tests.
"""
from __future__ import annotations

import numpy as np


def unpack(dims, p):
    p = np.asarray(p, np.float64)
    out, k = [], 0
    for i, o in zip(dims[:-1], dims[1:]):
        W = p[k:k + i * o].reshape(o, i)
        k += i * o
        b = p[k:k + o]
        k += o
        out.append((W, b))
    assert k == p.size
    return out


def forward(dims, p, X):
    h = np.asarray(X, np.float64)
    acts = [h]
    layers = unpack(dims, p)
    for l, (W, b) in enumerate(layers):
        z = h @ W.T + b
        h = np.maximum(z, 0.0) if l + 1 < len(layers) else z
        acts.append(h)
    return h, acts


def loss_and_grad(dims, p, X, t):
    y, acts = forward(dims, p, X)
    t = np.asarray(t, np.float64).reshape(y.shape)
    B = X.shape[0]
    loss = float(np.sum((y - t) ** 2) / B)
    d = 2.0 * (y - t) / B
    layers = unpack(dims, p)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        if l + 1 < len(layers):
            d = d * (acts[l + 1] > 0)
        grads[l] = (d.T @ acts[l], d.sum(axis=0))
        d = d @ W
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam(dims, p, X, t, epochs, lr, b1=0.9, b2=0.999, eps=1e-8):
    p = np.asarray(p, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    loss = None
    for step in range(1, epochs + 1):
        loss, g = loss_and_grad(dims, p, X, t)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p -= lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, loss


def finite_diff_grad(dims, p, X, t, h=1e-6):
    p = np.asarray(p, np.float64)
    g = np.zeros_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        g[k] = (loss_and_grad(dims, p + e, X, t)[0] - loss_and_grad(dims, p - e, X, t)[0]) / (2 * h)
    return g

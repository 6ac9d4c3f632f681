"""ORACLE - TEST INFRASTRUCTURE ONLY (checker, never the thing measured).

FP64 numpy restatement of the rank predictor's MLP as the library's kernels
define it (hmat_lib_amd/csrc/hip/hm_mlp.hip): nin -> H -> H -> 1, tanh
hidden layers, linear output, mean-squared-error loss, parameters as one
flat vector [W1 (nin x H, row i = input i), b1, W2 (H x H), b2, w3, b3].
The rank features follow hm_hip_rank_features (hm_compress.hip).

Pinned in tests/test_oracle.py: the gradient against central finite
differences of the loss (catches a dropped chain-rule term, a transposed
weight or a wrong activation derivative), the forward pass against a
direct evaluation of the formula, and the features against their
definition on hand-made spectra.  Not part of the reference (hmat_lib has
no learned component): the checker for this addition only.
"""
from __future__ import annotations

import numpy as np


def unpack(p, nin, nh):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + nin * nh].reshape(nin, nh); o += nin * nh
    b1 = p[o:o + nh]; o += nh
    W2 = p[o:o + nh * nh].reshape(nh, nh); o += nh * nh
    b2 = p[o:o + nh]; o += nh
    w3 = p[o:o + nh]; o += nh
    b3 = p[o]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, w3, b3


def nparams(nin, nh):
    return nin * nh + nh + nh * nh + nh + nh + 1


def forward(p, X, nh):
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1], nh)
    a1 = np.tanh(X @ W1 + b1)
    a2 = np.tanh(a1 @ W2 + b2)
    return a2 @ w3 + b3, a1, a2


def loss_and_grad(p, X, t, nh):
    """Mean squared error and its gradient (flat, same layout as p)."""
    X = np.asarray(X, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    B, nin = X.shape
    W1, b1, W2, b2, w3, b3 = unpack(p, nin, nh)
    y, a1, a2 = forward(p, X, nh)
    err = y - t
    loss = float(np.mean(err * err))
    dy = 2.0 * err / B
    gw3 = a2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, w3) * (1.0 - a2 * a2)
    gW2 = a1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (1.0 - a1 * a1)
    gW1 = X.T @ dz1
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def rank_features(extra, sig16, thr):
    """(nblock, n_extra + 17) model inputs: caller features, then
    log10(sigma_i / sigma_0) / 10 (i = 1..15, floored at 1e-20), the
    fraction of those ratios >= thr, log10(thr) / 10."""
    extra = np.asarray(extra, dtype=np.float64).reshape(len(sig16), -1)
    sig16 = np.asarray(sig16, dtype=np.float64)
    s0 = sig16[:, :1]
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(s0 > 0, sig16[:, 1:] / np.where(s0 > 0, s0, 1.0), 0.0)
    lr = np.log10(np.maximum(r, 1e-20)) / 10.0
    above = ((r >= thr) & (s0 > 0)).sum(1, keepdims=True) / 16.0
    th = np.full((len(sig16), 1), np.log10(thr) / 10.0)
    return np.hstack([extra, lr, above, th])

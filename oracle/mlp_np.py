"""FP64 NumPy restatement of the surrogate MLP's forward and backward -- TEST INFRASTRUCTURE.

Mirrors chemgen_amd/surrogate.py layer by layer: z_l = h_{l-1} W_l^T + b_l, h_l = SiLU(z_l)
(no activation on the output layer), loss = mean (zhat - z)^2.  With ``quantize`` the values
the GPU keeps in BF16 (inputs, weights, hidden activations, back-propagated gradients) are
rounded to BF16 (round to nearest even) at the same places, so the GPU gradients can be
compared with it; without it the function is plain FP64 calculus, which the tests pin by
finite differences.  Used only by tests as the checker.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(X, Ws, bs, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    h = q(X)
    acts, pres = [h], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ q(W).T + b
        pres.append(z)
        if l < len(Ws) - 1:
            h = q(silu(z))
            acts.append(h)
    return z, acts, pres


def features(P, T, Y, fscale, kin):
    """Network inputs (before BF16 rounding) of points P[n], T[n], Y[Ns, n]:
    (T - f0) f1, (ln P - f2) f3, (ln(Y_k + 1e-12) - f4) f5, (1000/T - f6) f7, zero padding to kin,
    evaluated in FP64 and rounded to FP32 (the GPU evaluates them in FP32, a few FP32 ulps
    away, before its BF16 conversion)."""
    f = np.asarray(fscale, dtype=np.float64)
    ns, n = Y.shape
    X = np.zeros((n, kin), dtype=np.float64)
    X[:, 0] = np.float32((T - f[0]) * f[1])
    X[:, 1] = np.float32((np.log(P) - f[2]) * f[3])
    X[:, 2:ns + 2] = np.float32((np.log(Y.T + 1e-12) - f[4]) * f[5])
    if kin > ns + 2:
        X[:, ns + 2] = np.float32((1000.0 / T - f[6]) * f[7])
    return X


def decode(z, s, aek):
    """wdot[Ns, n] = s_k sinh(z) and the element-conservation residual per point
    max_e |sum_k a_ek w_k| / sum_k |a_ek w_k| (0 where the denominator is 0)."""
    w = np.asarray(s, dtype=np.float64)[:, None] * np.sinh(np.asarray(z, dtype=np.float64).T)
    t = np.asarray(aek, dtype=np.float64)[:, :, None] * w[None, :, :]
    num, den = np.abs(t.sum(axis=1)), np.abs(t).sum(axis=1)
    r = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
    return w, r.max(axis=0)


def loss_and_grads(X, Z, Ws, bs, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    zh, acts, pres = forward(X, Ws, bs, quantize)
    n = zh.size
    loss = float(np.sum((zh - Z) ** 2) / n)
    G = q(2.0 * (zh - Z) / n)
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        gW[l] = G.T @ acts[l]
        gb[l] = G.sum(axis=0)
        if l > 0:
            dH = G @ q(Ws[l])
            G = q(dH * silu_grad(pres[l - 1]))
    return loss, gW, gb

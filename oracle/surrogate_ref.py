"""CPU oracle of the learned surrogate's network, FP64 numpy - TEST
INFRASTRUCTURE ONLY (the product never imports it; only tests/ and the
bench's checks do).

The reference has no learned part (README.md:31 of the reference says
learning belongs to another project), so this restates the network the
library defines (ice_crystal_phase_amd/csrc/surrogate.hip): inputs
[t, cos(2 pi k (v - 1/4)) (k=1..16), sin(pi f t), cos(pi f t) (f in
THETA_FREQ), c0..c3], four SiLU layers of
256, one log density per cell, softmax over the cells of f + log du_i, the
cross-entropy loss -sum P log p and the training objective's optional
total-variation term. Pinned by central finite differences
(tests/test_surrogate_oracle.py), which fix every term of the gradient.
"""
from __future__ import annotations

import numpy as np

H = 256
NF = 16
THETA_FREQ = np.array([1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96], np.float64)
NIN = 1 + NF + 2 * THETA_FREQ.size + 4


def u_edges(nt: int) -> np.ndarray:
    i = np.arange(nt + 1)
    k = np.minimum(i, nt - i)
    s = np.sin(0.5 * np.pi * k / nt) ** 2
    return np.where(i <= nt // 2, s, 1.0 - s)


def normalize_cond(cond) -> np.ndarray:
    a, lam, r, asp = (float(x) for x in cond)
    return np.array([a / 90.0, (lam - 550.0) / 150.0, (r - 1.5) / 1.5, np.log2(asp)])


def features(nt: int, np_: int, cond) -> np.ndarray:
    t = (np.arange(nt) + 0.5) / nt
    v = (np.arange(np_) + 0.5) / np_
    T, V = np.meshgrid(t, v, indexing="ij")
    k = np.arange(1, NF + 1)
    ang = 2 * np.pi * (V.reshape(-1, 1) - 0.25) * k
    tang = np.pi * T.reshape(-1, 1) * THETA_FREQ
    c = np.broadcast_to(normalize_cond(cond), (nt * np_, 4))
    return np.concatenate([T.reshape(-1, 1), np.cos(ang), np.sin(tang), np.cos(tang), c], axis=1)


def log_du(nt: int, np_: int) -> np.ndarray:
    return np.repeat(np.log(np.diff(u_edges(nt))), np_)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    o = 0
    W = []
    b = []
    for fan_in in (NIN, H, H, H):
        W.append(p[o:o + H * fan_in].reshape(H, fan_in))
        o += H * fan_in
        b.append(p[o:o + H])
        o += H
    w5 = p[o:o + H]
    b5 = p[o + H]
    assert o + H + 1 == p.size
    return W, b, w5, b5


def param_count() -> int:
    return H * NIN + H + 3 * (H * H + H) + H + 1


def silu(z):
    return z / (1 + np.exp(-z))


def forward(p, X, lw):
    W, b, w5, b5 = unpack(p)
    h = X
    zs, hs = [], [X]
    for l in range(4):
        z = h @ W[l].T + b[l]
        h = silu(z)
        zs.append(z)
        hs.append(h)
    f = h @ w5 + b5 + lw
    return f, zs, hs


def probs(p, nt, np_, cond) -> np.ndarray:
    f, _, _ = forward(p, features(nt, np_, cond), log_du(nt, np_))
    f = f - f.max()
    e = np.exp(f)
    return e / e.sum()


def loss_and_grad(p, nt, np_, cond, target, tv_weight: float = 0.0):
    """Cross entropy -sum P log softmax(f) and the gradient (packed layout) of
    the training objective cross entropy + tv_weight * total variation
    1/2 sum |q - P|, q = softmax(f): dF = q - P + tv_weight q (g - G) with
    g = sign(q - P) / 2 and G = sum g q (the subgradient of the total variation
    through the softmax)."""
    X = features(nt, np_, cond)
    lw = log_du(nt, np_)
    f, zs, hs = forward(p, X, lw)
    m = f.max()
    lse = m + np.log(np.exp(f - m).sum())
    P = np.asarray(target, np.float64).reshape(-1)
    loss = float(np.sum(P[P > 0] * (lse - f[P > 0])))
    q = np.exp(f - lse)
    dF = q - P
    if tv_weight:
        g = 0.5 * np.sign(q - P)
        dF = dF + tv_weight * q * (g - np.sum(g * q))
    W, b, w5, b5 = unpack(p)
    gW, gb = [None] * 4, [None] * 4
    gw5 = hs[4].T @ dF
    gb5 = dF.sum()
    dh = np.outer(dF, w5)
    for l in range(3, -1, -1):
        s = 1 / (1 + np.exp(-zs[l]))
        dz = dh * s * (1 + zs[l] * (1 - s))
        gW[l] = dz.T @ hs[l]
        gb[l] = dz.sum(axis=0)
        dh = dz @ W[l]
    out = []
    for l in range(4):
        out += [gW[l].reshape(-1), gb[l]]
    out += [gw5, np.array([gb5])]
    return loss, np.concatenate(out)


def objective(p, nt, np_, cond, target, tv_weight: float = 0.0) -> float:
    """Cross entropy + tv_weight * total variation of the network's cell
    probabilities against the target (the value loss_and_grad differentiates)."""
    loss, _ = loss_and_grad(p, nt, np_, cond, target)
    q = probs(p, nt, np_, cond)
    return loss + tv_weight * 0.5 * float(np.abs(q - np.asarray(target, np.float64).reshape(-1)).sum())

"""CPU pins of the surrogate's FP64 oracle (oracle/surrogate_ref.py): central
finite differences fix every gradient term; the parameter count and packed
layout agree with the library's C-ABI (no device call)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "oracle"))
import surrogate_ref as ref  # noqa: E402

import ice_crystal_phase_amd as ip  # noqa: E402


def _params(seed: int, scale: float = 1.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, ref.param_count())
    p[: 256 * ref.NIN] /= np.sqrt(ref.NIN)
    p[256 * ref.NIN:] /= 16.0
    return p * scale


def test_param_count_matches_library():
    assert ip.lib().iph_sur_param_count() == ref.param_count() == 256 * 47 + 256 + 3 * 65792 + 257


def test_features_and_weights():
    X = ref.features(6, 8, (30.0, 550.0, 1.5, 1.0))
    assert X.shape == (48, 47)
    t = (np.arange(6) + 0.5) / 6
    assert np.allclose(X[:, 0].reshape(6, 8)[:, 0], t)
    # phi features depend on the column only, theta features on the row only
    v = (np.arange(8) + 0.5) / 8
    assert np.allclose(X[:8, 1 + 2], np.cos(2 * np.pi * 3 * (v - 0.25)))
    assert np.allclose(X[::8, 17 + 4], np.sin(np.pi * 6 * t)) and np.allclose(X[::8, 30 + 12], np.cos(np.pi * 96 * t))
    assert np.allclose(X[:, 43:], [30 / 90, 0.0, 0.0, 0.0])
    # the phi features are even about v = 1/4: columns j and np/2 - 1 - j (mod np) match
    Xr = X.reshape(6, 8, 47)
    for j in range(8):
        assert np.array_equal(Xr[:, j, 1:17], Xr[:, (8 // 2 - 1 - j) % 8, 1:17]) or np.allclose(
            Xr[:, j, 1:17], Xr[:, (8 // 2 - 1 - j) % 8, 1:17], rtol=0, atol=1e-12)
    lw = ref.log_du(6, 8)
    assert np.isclose(np.exp(lw).reshape(6, 8)[:, 0].sum(), 1.0)


def test_gradient_matches_finite_differences():
    nt, np_ = 6, 8
    rng = np.random.default_rng(7)
    target = rng.random(nt * np_) ** 3
    target /= target.sum()
    cond = (-40.0, 470.0, 2.2, 0.83)
    p = _params(1)
    loss, g = ref.loss_and_grad(p, nt, np_, cond, target)
    assert np.isfinite(loss) and g.shape == p.shape
    n1 = 256 * ref.NIN
    idx = np.concatenate([rng.choice(n1, 12, replace=False),                           # W1
                          np.arange(43, n1, ref.NIN)[rng.choice(256, 4, replace=False)],  # W1, condition column
                          n1 + rng.choice(256, 4, replace=False),                      # b1
                          n1 + 256 + rng.choice(65536 + 256, 12, replace=False),       # layer 2
                          n1 + 256 + 2 * 65792 + rng.choice(65792, 8, replace=False),  # layer 4
                          ref.param_count() - 257 + rng.choice(257, 6, replace=False)])  # w5, b5
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (ref.loss_and_grad(pp, nt, np_, cond, target)[0] - ref.loss_and_grad(pm, nt, np_, cond, target)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_tv_gradient_matches_finite_differences():
    """The gradient of cross entropy + w * total variation: central differences
    of the objective (the total variation is piecewise linear in q, smooth away
    from q = P, which a random target keeps far from every cell)."""
    nt, np_ = 6, 8
    rng = np.random.default_rng(9)
    target = rng.random(nt * np_) ** 3
    target /= target.sum()
    cond = (25.0, 640.0, 0.7, 1.0)
    p = _params(3)
    q = ref.probs(p, nt, np_, cond)
    assert np.min(np.abs(q - target)) > 1e-5
    for w in (1.0, 7.5):
        _, g = ref.loss_and_grad(p, nt, np_, cond, target, tv_weight=w)
        _, g0 = ref.loss_and_grad(p, nt, np_, cond, target)
        assert np.linalg.norm(g - g0) > 1e-3 * np.linalg.norm(g0)  # the term contributes
        n1 = 256 * ref.NIN
        idx = np.concatenate([rng.choice(n1, 8, replace=False), n1 + 256 + rng.choice(65536 + 256, 8, replace=False),
                              ref.param_count() - 257 + rng.choice(257, 6, replace=False)])
        h = 1e-6
        for i in idx:
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (ref.objective(pp, nt, np_, cond, target, w) - ref.objective(pm, nt, np_, cond, target, w)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (w, i, fd, g[i])


def test_tv_gradient_is_shift_invariant():
    """Closed form: the softmax is invariant to a common shift of the logits, so
    the objective's derivative along the output bias is sum_k dF_k = 1 - sum P
    (= 0 for a normalized target) with or without the total-variation term,
    whose part sum_k q_k (g_k - G) = G - G vanishes only if G is subtracted."""
    nt, np_ = 5, 6
    rng = np.random.default_rng(4)
    target = rng.random(nt * np_)
    target /= target.sum()
    p = _params(4)
    cond = (0.0, 550.0, 1.0, 1.0)
    _, g = ref.loss_and_grad(p, nt, np_, cond, target, tv_weight=3.0)
    assert abs(g[-1]) < 1e-12
    q = ref.probs(p, nt, np_, cond)
    sg = 0.5 * np.sign(q - target)
    assert abs(np.sum(q * sg)) > 1e-3  # the subtracted G is not trivially zero here


def test_probabilities_normalized_and_loss_is_cross_entropy():
    nt, np_ = 5, 7
    p = _params(2)
    q = ref.probs(p, nt, np_, (10.0, 600.0, 1.0, 1.0))
    assert np.isclose(q.sum(), 1.0) and np.all(q > 0)
    loss, _ = ref.loss_and_grad(p, nt, np_, (10.0, 600.0, 1.0, 1.0), q)
    assert np.isclose(loss, -np.sum(q * np.log(q)))

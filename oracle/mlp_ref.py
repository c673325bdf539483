"""FP64 numpy restatement of the learned LES closure and its training step.

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker). The closure has no
reference counterpart: it replaces the Smagorinsky eddy viscosity of
get_Smagorinsky_LES_tau (src/lbm/kernels/steps/collide.mojo:70-73). MLP 9 -> 32 -> 32 -> 1,
SiLU hidden units, softplus output times nu_scale; loss = mean (softplus(out) - y/nu_scale)^2;
Adam with bias correction (Kingma & Ba 2015). Pinned in tests/test_oracle_mlp.py by central
finite differences of the loss and the closed form of Adam's first step.
"""
from __future__ import annotations

import numpy as np

IN, H = 9, 32
W1, B1 = 0, 288
W2, B2 = 320, 1344
W3, B3 = 1376, 1408
NPARAM = 1409
MEAN, ISTD, LO, HI, NUS = 1409, 1418, 1427, 1436, 1445
TOTAL = 1446


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[W1:B1].reshape(IN, H), P[B1:W2], P[W2:B2].reshape(H, H), P[B2:W3], P[W3:B3], P[B3],
            P[MEAN:ISTD], P[ISTD:LO], P[LO:HI], P[HI:NUS], P[NUS])


def _silu(a):
    return a / (1.0 + np.exp(-a))


def _sig(a):
    return 1.0 / (1.0 + np.exp(-a))


def _softplus(a):
    return np.where(a > 20.0, a, np.log1p(np.exp(np.minimum(a, 20.0))))


def forward(P, X):
    """nu_t = softplus(MLP(z)) * nu_scale for raw features X [n, 9]; also returns the cache."""
    w1, b1, w2, b2, w3, b3, mean, istd, _, _, nus = unpack(P)
    z = (np.asarray(X, np.float64) - mean) * istd
    a1 = z @ w1 + b1
    h1 = _silu(a1)
    a2 = h1 @ w2 + b2
    h2 = _silu(a2)
    out = h2 @ w3 + b3
    return _softplus(out) * nus, (z, a1, h1, a2, h2, out)


def loss_grad(P, X, y):
    """(loss, d loss / d weights [1409]) for loss = mean (softplus(out) - y / nu_scale)^2."""
    w1, b1, w2, b2, w3, b3, mean, istd, _, _, nus = unpack(P)
    _, (z, a1, h1, a2, h2, out) = forward(P, X)
    n = len(out)
    r = _softplus(out) - np.asarray(y, np.float64) / nus
    loss = float(np.mean(r * r))
    go = 2.0 * r / n * _sig(out)
    g2 = go[:, None] * w3[None, :] * (_sig(a2) * (1.0 + a2 * (1.0 - _sig(a2))))
    gh1 = g2 @ w2.T
    g1 = gh1 * (_sig(a1) * (1.0 + a1 * (1.0 - _sig(a1))))
    grad = np.zeros(NPARAM)
    grad[W1:B1] = (z.T @ g1).ravel()
    grad[B1:W2] = g1.sum(0)
    grad[W2:B2] = (h1.T @ g2).ravel()
    grad[B2:W3] = g2.sum(0)
    grad[W3:B3] = h2.T @ go
    grad[B3] = go.sum()
    return loss, grad


def adam(P, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on the weights; returns new (P, m, v)."""
    P = np.array(P, np.float64)
    m = b1 * np.asarray(m, np.float64) + (1 - b1) * g
    v = b2 * np.asarray(v, np.float64) + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    P[:NPARAM] -= lr * mh / (np.sqrt(vh) + eps)
    return P, m, v


def closure_nu(P, x, nu_smag):
    """The collide rule: the MLP's nu_t where every raw feature lies in [lo, hi], else nu_smag."""
    _, _, _, _, _, _, _, _, lo, hi, _ = unpack(P)
    x = np.asarray(x, np.float32).astype(np.float64)
    inside = np.all((x >= lo) & (x <= hi), axis=-1)
    nu, _ = forward(P, x)
    return np.where(inside, nu, nu_smag)


def random_params(seed=0, nu_scale=1e-3, feature_scale=None):
    """Seeded parameter block with fan-in scaled weights and a wide training range."""
    rng = np.random.default_rng(seed)
    P = np.zeros(TOTAL)
    P[W1:B1] = rng.standard_normal(IN * H) / np.sqrt(IN)
    P[B1:W2] = 0.1 * rng.standard_normal(H)
    P[W2:B2] = rng.standard_normal(H * H) / np.sqrt(H)
    P[B2:W3] = 0.1 * rng.standard_normal(H)
    P[W3:B3] = rng.standard_normal(H) / np.sqrt(H)
    P[B3] = -1.0
    P[MEAN:ISTD] = 0.0
    P[ISTD:LO] = 1.0 if feature_scale is None else 1.0 / np.asarray(feature_scale, np.float64)
    P[LO:HI] = -3e38
    P[HI:NUS] = 3e38
    P[NUS] = nu_scale
    return P.astype(np.float32)

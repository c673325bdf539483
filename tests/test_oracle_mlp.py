"""Pins the FP64 closure restatement (oracle/mlp_ref.py): its analytic gradient against
central finite differences of the loss, and Adam against the closed form of its first
step (m_hat = g, v_hat = g^2, so the update is -lr g / (|g| + eps))."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def _data(n=40, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 9))
    y = np.abs(rng.standard_normal(n)) * 1e-3
    return X, y


def test_gradient_matches_finite_differences():
    P = M.random_params(3).astype(np.float64)
    X, y = _data()
    _, g = M.loss_grad(P, X, y)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(M.NPARAM, 60, replace=False), [M.B3, M.B1, M.W3, M.B2]])
    for i in idx:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (M.loss_grad(Pp, X, y)[0] - M.loss_grad(Pm, X, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_gradient_wrong_sign_would_fail():
    # negative control: the finite-difference check distinguishes g from -g
    P = M.random_params(4).astype(np.float64)
    X, y = _data()
    _, g = M.loss_grad(P, X, y)
    i = int(np.argmax(np.abs(g)))
    h = 1e-6
    Pp, Pm = P.copy(), P.copy()
    Pp[i] += h
    Pm[i] -= h
    fd = (M.loss_grad(Pp, X, y)[0] - M.loss_grad(Pm, X, y)[0]) / (2 * h)
    assert abs(fd + g[i]) > 1e-3 * abs(g[i])


def test_adam_first_step_closed_form():
    P = M.random_params(5).astype(np.float64)
    g = np.random.default_rng(2).standard_normal(M.NPARAM)
    lr, eps = 1e-3, 1e-8
    P1, m, v = M.adam(P, np.zeros(M.NPARAM), np.zeros(M.NPARAM), g, 1, lr, eps=eps)
    assert np.allclose(P1[:M.NPARAM] - P[:M.NPARAM], -lr * g / (np.abs(g) + eps), rtol=1e-12, atol=1e-15)
    assert np.array_equal(P1[M.NPARAM:], P[M.NPARAM:])


def test_closure_range_fallback():
    P = M.random_params(6)
    P[M.LO:M.HI] = -1.0
    P[M.HI:M.NUS] = 1.0
    x = np.zeros((3, 9))
    x[1, 4] = 2.0  # outside
    nu = M.closure_nu(P, x, np.array([7.0, 7.0, 7.0]))
    assert nu[1] == 7.0 and nu[0] != 7.0 and nu[2] != 7.0


@pytest.mark.parametrize("a", [-30.0, -1.0, 0.0, 0.5, 25.0])
def test_softplus_and_silu_values(a):
    assert M._softplus(np.array(a)) == pytest.approx(np.log1p(np.exp(a)) if a <= 20 else a, rel=1e-12)
    assert M._silu(np.array(a)) == pytest.approx(a / (1 + np.exp(-a)), rel=1e-12)


def test_oracle_closure_matches_fp64_mlp():
    """The C++ oracle's closure (the one its collide uses: bf16-rounded inputs, weights of the two hidden
    layers and first hidden activations, as the matrix cores see them) agrees with the FP64 MLP to bf16
    accuracy; a transposed weight matrix or a wrong activation would miss by O(1)."""
    from oracle import oracle as O

    P = M.random_params(11, nu_scale=1.0)
    rng = np.random.default_rng(12)
    X = rng.standard_normal((400, 9)).astype(np.float32)
    got = O.mlp_closure(P, X)
    ref = M.closure_nu(P, X, np.full(len(X), -1.0))
    assert np.all(got > 0)
    np.testing.assert_allclose(got, ref, rtol=3e-2)
    assert np.median(np.abs(got / ref - 1)) < 5e-3
    P2 = P.copy()
    P2[M.HI:M.NUS] = 0.5  # most samples leave the training range
    out = O.mlp_closure(P2, X)
    inside = np.all(X <= 0.5, axis=1)
    assert np.all(out[~inside] == -1.0) and np.all(out[inside] > 0)

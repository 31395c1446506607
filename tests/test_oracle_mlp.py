"""Pins of the MLP oracle (the checker of the cost model's HIP kernels):
the analytic gradient equals central finite differences of the loss, the
masked loss equals its definition, and a planted mistake in the gradient
(GELU's derivative without its second term) is caught."""
import numpy as np

import oracle.mlp as om

N_IN, HID, OUT = 5, 7, 3


def case(seed=0, n=11):
    rng = np.random.default_rng(seed)
    k = sum(int(np.prod(s)) for _, s in om.sizes(N_IN, HID, OUT))
    params = rng.standard_normal(k) * 0.7
    X = rng.standard_normal((n, N_IN))
    Y = rng.standard_normal((n, OUT))
    M = (rng.uniform(size=(n, OUT)) > 0.3).astype(np.float64)
    return params, X, Y, M, 1.0 / M.sum()


def fd_grad(params, X, Y, M, ic, h=1e-6):
    g = np.zeros_like(params)
    for k in range(params.size):
        e = np.zeros_like(params)
        e[k] = h
        g[k] = (om.loss(params + e, X, Y, M, ic, N_IN, HID, OUT)
                - om.loss(params - e, X, Y, M, ic, N_IN, HID, OUT)) / (2 * h)
    return g


def test_gradient_equals_finite_differences():
    for seed in range(3):
        params, X, Y, M, ic = case(seed)
        L, g = om.loss_and_grad(params, X, Y, M, ic, N_IN, HID, OUT)
        assert np.isclose(L, om.loss(params, X, Y, M, ic, N_IN, HID, OUT))
        np.testing.assert_allclose(g, fd_grad(params, X, Y, M, ic), rtol=1e-6, atol=1e-8)


def test_masked_loss_definition():
    params, X, Y, M, ic = case(4)
    P, _ = om.forward(params, X, N_IN, HID, OUT)
    want = sum((P[r, t] - Y[r, t]) ** 2 for r in range(X.shape[0]) for t in range(OUT)
               if M[r, t] > 0) / M.sum()
    assert np.isclose(om.loss(params, X, Y, M, ic, N_IN, HID, OUT), want)
    # masked targets do not matter
    Y2 = np.where(M > 0, Y, 1e6)
    assert np.isclose(om.loss(params, X, Y2, M, ic, N_IN, HID, OUT), want)


def test_gelu_closed_forms():
    assert om.gelu(0.0) == 0.0
    assert np.isclose(om.gelu(1.0), 0.8413447460685429)  # 1 * Phi(1)
    assert np.isclose(om.gelu_grad(0.0), 0.5)


def test_planted_gradient_mistake_is_caught():
    params, X, Y, M, ic = case(5)
    fd = fd_grad(params, X, Y, M, ic)
    good = om.gelu_grad
    try:
        om.gelu_grad = lambda z: 0.5 * (1.0 + om.erf(z / om.SQRT2))  # drops z phi(z)
        _, g_bad = om.loss_and_grad(params, X, Y, M, ic, N_IN, HID, OUT)
    finally:
        om.gelu_grad = good
    assert not np.allclose(g_bad, fd, rtol=1e-4, atol=1e-6)


def test_adam_first_step_moves_by_lr():
    """Adam's first bias-corrected step is lr * sign(g) (|g| >> eps)."""
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = om.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-4)

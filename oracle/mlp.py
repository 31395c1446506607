"""CPU restatement (float64) of the learned dispatcher's MLP -- TEST
INFRASTRUCTURE ONLY: the checker for the HIP forward/backward/Adam kernels.

No reference counterpart (the reference has no learned dispatch).  The math
is the standard one: Z1 = X W1^T + b1, A1 = gelu(Z1), Z2 = A1 W2^T + b2,
A2 = gelu(Z2), P = A2 W3^T + b3, gelu(z) = z Phi(z) with the exact normal
CDF, loss = sum M (P - Y)^2 * inv_count.  Parameters are flat in the order
W1 b1 W2 b2 W3 b3, W_l[out][in].  Pinned by central finite differences of the
loss (tests/test_oracle_mlp.py), which do not use the analytic backward here.
"""
import numpy as np
from scipy.special import erf

SQRT2 = np.sqrt(2.0)


def sizes(n_in, hid, out):
    return [("W1", (hid, n_in)), ("b1", (hid,)), ("W2", (hid, hid)), ("b2", (hid,)),
            ("W3", (out, hid)), ("b3", (out,))]


def unpack(params, n_in, hid, out):
    params = np.asarray(params, dtype=np.float64)
    res, off = {}, 0
    for name, shp in sizes(n_in, hid, out):
        k = int(np.prod(shp))
        res[name] = params[off:off + k].reshape(shp)
        off += k
    assert off == params.size
    return res


def pack(p, n_in, hid, out):
    return np.concatenate([p[name].ravel() for name, _ in sizes(n_in, hid, out)])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / SQRT2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / SQRT2)) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, X, n_in, hid, out):
    p = unpack(params, n_in, hid, out)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ p["W1"].T + p["b1"]
    A1 = gelu(Z1)
    Z2 = A1 @ p["W2"].T + p["b2"]
    A2 = gelu(Z2)
    P = A2 @ p["W3"].T + p["b3"]
    return P, (X, Z1, A1, Z2, A2)


def loss(params, X, Y, M, inv_count, n_in, hid, out):
    P, _ = forward(params, X, n_in, hid, out)
    M = np.asarray(M, dtype=np.float64)
    D = np.where(M > 0, P - Y, 0.0)
    return float(np.sum(D * D) * inv_count)


def loss_and_grad(params, X, Y, M, inv_count, n_in, hid, out):
    p = unpack(params, n_in, hid, out)
    P, (X, Z1, A1, Z2, A2) = forward(params, X, n_in, hid, out)
    M = np.asarray(M, dtype=np.float64)
    D = np.where(M > 0, P - np.asarray(Y, dtype=np.float64), 0.0)
    L = float(np.sum(D * D) * inv_count)
    dP = 2.0 * D * inv_count
    g = {"W3": dP.T @ A2, "b3": dP.sum(0)}
    dZ2 = (dP @ p["W3"]) * gelu_grad(Z2)
    g["W2"], g["b2"] = dZ2.T @ A1, dZ2.sum(0)
    dZ1 = (dZ2 @ p["W2"]) * gelu_grad(Z1)
    g["W1"], g["b1"] = dZ1.T @ X, dZ1.sum(0)
    return L, pack(g, n_in, hid, out)


def adam_step(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v

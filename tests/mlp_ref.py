"""numpy restatement of the step-scale model's forward and backward passes (used by the tests).

With rounded=True it applies the kernels' BF16 roundings: inputs, weights, the first hidden layer,
and the dz operands of the weight-gradient products; everything else is FP64.
"""
import numpy as np

NI, NH = 32, 64
O_W1, O_B1 = 0, NI * NH
O_W2 = O_B1 + NH
O_B2 = O_W2 + NH * NH
O_W3 = O_B2 + NH
O_B3 = O_W3 + NH


def bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[O_W1:O_B1].reshape(NI, NH), p[O_B1:O_W2], p[O_W2:O_B2].reshape(NH, NH), p[O_B2:O_W3], p[O_W3:O_B3], p[O_B3])


def forward_ref(p, x, rounded=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if rounded else (lambda a: np.asarray(a, np.float64))
    z1 = r(x) @ r(W1) + b1
    h1 = r(silu(z1).astype(np.float32)) if rounded else silu(z1)
    z2 = h1 @ r(W2) + b2
    z3 = silu(z2) @ w3 + b3
    s = 4.0 / (1.0 + np.exp(-z3))
    return s, (z1, h1, z2, z3)


def grad_ref(p, x, t, rounded=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if rounded else (lambda a: np.asarray(a, np.float64))
    s, (z1, h1, z2, z3) = forward_ref(p, x, rounded)
    dz3 = (s - t) * s * (1.0 - 0.25 * s)
    dz2 = dz3[:, None] * w3[None, :] * silu_grad(z2)
    dh1 = r(dz2.astype(np.float32)) @ r(W2).T
    dz1 = dh1 * silu_grad(z1)
    g = np.zeros(O_B3 + 1)
    g[O_W1:O_B1] = (r(x).T @ r(dz1.astype(np.float32))).ravel()
    g[O_B1:O_W2] = dz1.sum(0)
    g[O_W2:O_B2] = (h1.T @ r(dz2.astype(np.float32))).ravel()
    g[O_B2:O_W3] = dz2.sum(0)
    g[O_W3:O_B3] = (dz3[:, None] * silu(z2)).sum(0)
    g[O_B3] = dz3.sum()
    return g, 0.5 * np.sum((s - t) ** 2)


def sample_problem(n=200, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, NI)).astype(np.float32)
    t = rng.uniform(0.2, 3.8, n).astype(np.float32)
    return x, t



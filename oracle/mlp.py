"""FP64 numpy oracle of the cost-model MLP -- TEST INFRASTRUCTURE ONLY.

The reference has no cost model (its scheduling is a static split,
python-bh/barneshut/implementations/parla.py:188,248), so this oracle follows
the layer definition the library documents (include/bh_hip.h): an MLP
8 -> 32 -> 32 -> 1 with ReLU, parameters packed as W1[32x8], b1[32],
W2[32x32], b2[32], W3[32], b3[1] (row-major, out x in), loss
mean((y - t)^2), Adam(beta1 0.9, beta2 0.999, eps 1e-8).
Pinned in tests by finite differences (backward) and hand-computed cases
(forward).
"""
import numpy as np

NF, NH = 8, 32
OFF_W1, OFF_B1 = 0, NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NF), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(NH, NH), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3 + b3


def loss(p, X, t):
    e = forward(p, X) - np.asarray(t, np.float64)
    return float(np.mean(e * e))


def backward(p, X, t):
    """Gradient of mean((y - t)^2) w.r.t. the packed parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ W3 + b3
    dy = 2.0 * (y - np.asarray(t, np.float64)) / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = dy @ h2
    g[OFF_B3] = dy.sum()
    d2 = (dy[:, None] * W3[None, :]) * (a2 > 0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    g[OFF_W1:OFF_B1] = (d1.T @ X).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    return g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v

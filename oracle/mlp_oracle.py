"""ORACLE -- test infrastructure only.

FP64 numpy restatement of the step-size proposal network of the device library
(8 -> 64 -> 64 -> 1, GELU hidden layers, loss mean((y - t_n)^2) on normalised
targets): forward pass, loss and the gradient with respect to every parameter,
by hand-written backpropagation.  The network has no counterpart in the
reference; the gradient formula is pinned by central finite differences of the
loss (tests/test_mlp_oracle.py), and the device kernels are checked against
this restatement evaluated at their BF16-rounded weights.
"""
import numpy as np
from scipy.special import erf

NIN, NH = 8, 64
O_W1 = 0
O_B1 = O_W1 + NH * NIN
O_W2 = O_B1 + NH
O_B2 = O_W2 + NH * NH
O_W3 = O_B2 + NH
O_B3 = O_W3 + NH
NP = O_B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[O_W1:O_B1].reshape(NH, NIN), p[O_B1:O_W2], p[O_W2:O_B2].reshape(NH, NH), p[O_B2:O_W3],
            p[O_W3:O_B3], p[O_B3])


def bf16(p):
    """FP32 values rounded to BF16 (nearest even), returned as float64."""
    u = np.asarray(p, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def normalise(norm, X):
    norm = np.asarray(norm, np.float64)
    return (np.asarray(X, np.float64) - norm[:NIN]) / norm[NIN:2 * NIN]


def forward(p, norm, X):
    """Normalised outputs y (n,) and the cache for backpropagation."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    xn = normalise(norm, X)
    z1 = xn @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    y = h2 @ w3 + b3
    return y, (xn, z1, h1, z2, h2)


def loss(p, norm, X, T):
    norm = np.asarray(norm, np.float64)
    y, _ = forward(p, norm, X)
    t = (np.asarray(T, np.float64) - norm[2 * NIN]) / norm[2 * NIN + 1]
    return float(np.mean((y - t) ** 2))


def loss_grad(p, norm, X, T):
    norm = np.asarray(norm, np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (xn, z1, h1, z2, h2) = forward(p, norm, X)
    t = (np.asarray(T, np.float64) - norm[2 * NIN]) / norm[2 * NIN + 1]
    n = y.shape[0]
    e = y - t
    g = 2.0 * e / n                       # d loss / d y
    d2 = (g[:, None] * w3[None, :]) * gelu_d(z2)
    d1 = (d2 @ W2) * gelu_d(z1)
    grad = np.zeros(NP)
    grad[O_W1:O_B1] = (d1.T @ xn).reshape(-1)
    grad[O_B1:O_W2] = d1.sum(0)
    grad[O_W2:O_B2] = (d2.T @ h1).reshape(-1)
    grad[O_B2:O_W3] = d2.sum(0)
    grad[O_W3:O_B3] = h2.T @ g
    grad[O_B3] = g.sum()
    return float(np.mean(e * e)), grad

"""CPU restatement of the scheduler's cost-model MLP (test infrastructure).

Parameter layout: W1[32][8], b1[32], W2[32][32], b2[32], W3[32], b3; ReLU
hidden layers, linear output; loss 0.5 * sum (f(x) - t)^2.  float64 so it
can serve as the reference the float32 device kernels are checked against.
"""
import numpy as np

NI, NH = 8, 32
N_PARAMS = NH * NI + NH + NH * NH + NH + NH + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + NH * NI].reshape(NH, NI); o += NH * NI
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64).reshape(-1, NI)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3 + b3, (X, h1, h2)


def gradient(p, X, T):
    """Summed gradient and summed loss of 0.5 (f - t)^2 (backpropagation)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (X, h1, h2) = forward(p, X)
    e = y - np.asarray(T, np.float64)
    gW3 = h2.T @ e
    gb3 = e.sum()
    d2 = np.outer(e, W3) * (h2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (h1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return g, 0.5 * float((e * e).sum())

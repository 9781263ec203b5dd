"""FP64 numpy restatement of the learned-guess network (TEST INFRASTRUCTURE ONLY): forward pass and the gradient of
mean((y - t)^2), checked in tests against finite differences and used as the oracle of the HIP training kernels."""
import numpy as np
from math import erf, sqrt, pi

N_IN, N_H = 175, 128
_erf = np.vectorize(erf)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    o = 0
    W1 = P[o:o + N_IN * N_H].reshape(N_IN, N_H); o += N_IN * N_H
    b1 = P[o:o + N_H]; o += N_H
    W2 = P[o:o + N_H * N_H].reshape(N_H, N_H); o += N_H * N_H
    b2 = P[o:o + N_H]; o += N_H
    w3 = P[o:o + N_H]; o += N_H
    b3 = P[o]
    return W1, b1, W2, b2, w3, b3


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / sqrt(2.0)))


def dgelu(z):
    return 0.5 * (1.0 + _erf(z / sqrt(2.0))) + z * np.exp(-0.5 * z * z) / sqrt(2.0 * pi)


def forward(P, X):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    Z1 = X @ W1 + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2 + b2
    H2 = gelu(Z2)
    return H2 @ w3 + b3, (Z1, H1, Z2, H2)


def loss_grad(P, X, T):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    y, (Z1, H1, Z2, H2) = forward(P, X)
    B = X.shape[0]
    d = y - T
    loss = float(np.mean(d * d))
    dy = 2.0 * d / B
    gw3 = H2.T @ dy
    gb3 = dy.sum()
    dZ2 = np.outer(dy, w3) * dgelu(Z2)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * dgelu(Z1)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])

"""FP64 numpy restatement of the device MLP (F -> 128 -> 128 -> 1, tanh): forward pass,
mean-squared-error loss and its gradient.  TEST INFRASTRUCTURE ONLY (the checker of the
GPU training kernels); pinned by central finite differences in tests/test_mlp_oracle.py.
Parameter layout as the device's: W1[F][H] b1[H] W2[H][H] b2[H] W3[H] b3."""
import numpy as np

H = 128


def unpack(p, F):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, X):
    F = X.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(p, F)
    h1 = np.tanh(X @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    return h2 @ W3 + b3, (h1, h2)


def loss_grad(p, X, y):
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    n, F = X.shape
    W1, b1, W2, b2, W3, b3 = unpack(p, F)
    pred, (h1, h2) = forward(p, X)
    r = pred - y
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, W3) * (1.0 - h2 * h2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2.T) * (1.0 - h1 * h1)
    gW1 = X.T @ dz1
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g

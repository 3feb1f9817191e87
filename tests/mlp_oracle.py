"""CPU restatement of the learned indicator's network (test infrastructure only).

Float64 forward, loss and analytic backward of the 112 -> 64 -> 64 -> 4 SiLU
network with the same weight layout, loss (class-weighted softmax cross
entropy on outputs 0..2 plus logistic loss on output 3, both averaged over the
batch) and conventions as the device kernels.  Pinned by central finite
differences in tests/test_mlp.py.
"""
import numpy as np

K, H, NO = 112, 64, 4


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    W1 = w[o:o + K * H].reshape(K, H); o += K * H
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    W3 = w[o:o + H * NO].reshape(H, NO); o += H * NO
    b3 = w[o:o + NO]; o += NO
    return W1, b1, W2, b2, W3, b3


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(w, X):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    z1 = X @ W1 + b1; h1 = z1 * sig(z1)
    z2 = h1 @ W2 + b2; h2 = z2 * sig(z2)
    return h2 @ W3 + b3, (X, z1, h1, z2, h2)


def loss_and_grad(w, X, cls, smooth, cw=(1.0, 1.0, 1.0)):
    X = np.asarray(X, dtype=np.float64)
    B = X.shape[0]
    W1, b1, W2, b2, W3, b3 = unpack(w)
    o, (X, z1, h1, z2, h2) = forward(w, X)
    lg = o[:, :3]
    mx = lg.max(1, keepdims=True)
    e = np.exp(lg - mx)
    p = e / e.sum(1, keepdims=True)
    wy = np.asarray(cw)[cls]
    s = sig(o[:, 3])
    t = np.asarray(smooth, dtype=np.float64)
    loss = (-wy * np.log(np.maximum(p[np.arange(B), cls], 1e-300))
            - (t * np.log(np.maximum(s, 1e-300)) + (1 - t) * np.log(np.maximum(1 - s, 1e-300)))).mean()
    d3 = np.zeros_like(o)
    oh = np.zeros_like(p); oh[np.arange(B), cls] = 1
    d3[:, :3] = wy[:, None] * (p - oh) / B
    d3[:, 3] = (s - t) / B
    gW3 = h2.T @ d3; gb3 = d3.sum(0)
    dsil = lambda z: sig(z) + z * sig(z) * (1 - sig(z))
    d2 = (d3 @ W3.T) * dsil(z2)
    gW2 = h1.T @ d2; gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * dsil(z1)
    gW1 = X.T @ d1; gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g

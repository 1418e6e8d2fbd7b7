"""CPU restatement (float64 numpy) of the field surrogate's network, its
loss and gradient, and of its output decoding.  Test infrastructure: the
device training and inference kernels are checked against it, and it is
pinned by finite differences (tests/test_mlp_ref.py)."""
import numpy as np

H = 64
NPARAM = 4546
K0, K1 = 0.7978845608028654, 0.044715
LN_RMIN = np.log(0.02)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[0:192].reshape(64, 3), p[192:256], p[256:4352].reshape(64, 64), p[4352:4416],
            p[4416:4544].reshape(2, 64), p[4544:4546])


def gelu(z):
    a = K0 * z * (1 + K1 * z * z)
    t = np.tanh(a)
    g = 0.5 * z * (1 + t)
    dg = 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * K0 * (1 + 3 * K1 * z * z)
    return g, dg


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = X @ W1.T + b1
    h1, dg1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2, dg2 = gelu(z2)
    return h2 @ W3.T + b3, (X, z1, h1, dg1, z2, h2, dg2)


def loss_grad(p, recs, norm):
    """Mean over the records of (o0 - y0)^2 + (o1 - y1)^2, y = normalized targets."""
    recs = np.asarray(recs, dtype=np.float64)
    X = recs[:, :3]
    Y = np.stack([(recs[:, 3] - norm[0]) / norm[1], (recs[:, 4] - norm[2]) / norm[3]], 1)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    out, (X, z1, h1, dg1, z2, h2, dg2) = forward(p, X)
    n = len(X)
    e = out - Y
    loss = np.sum(e * e) / n
    dout = 2 * e / n
    gW3 = dout.T @ h2
    gb3 = dout.sum(0)
    dz2 = (dout @ W3) * dg2
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * dg1
    gW1 = dz1.T @ X
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def features(au, av, r):
    return np.stack([au / 3 - 1, av / 3 - 1, (np.log(r) - LN_RMIN / 2) / (-LN_RMIN / 2)], -1)


def envelope(au, av, r):
    rho2 = au * au + (av * r) ** 2
    qx = rho2 / (1 + r)
    fx = np.where(qx < 1e-12, 1.0, -np.expm1(-qx) / np.maximum(qx, 1e-300))
    return -fx / (1 + r), -np.ones_like(fx)


def decode(au, av, r, out, norm):
    bx, by = envelope(au, av, r)
    gx = au * bx * np.exp(norm[0] + norm[1] * out[:, 0])
    gy = av * by * np.exp(norm[2] + norm[3] * out[:, 1])
    return gx, gy

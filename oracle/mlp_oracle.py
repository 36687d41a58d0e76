"""FP64 CPU restatement of the learned-proposal network (TEST INFRASTRUCTURE
ONLY: used by tests/ as the checker of the GPU inference and training
kernels).

Network: inputs f (n_env of +-1, zero padded to 48), h = W1 f + b1,
z = W2 . relu(h) + b2; the flipped-occupation output uses -f.  A sample with
logit scale t (beta * 0.1 eV) has logit t * z.
Loss: mean binary cross-entropy of sigmoid(t z) against soft labels y.
Parameter vector: [W1 (64 x 48 row-major) | b1 (64) | W2 (64) | b2].
"""
import numpy as np

H, K = 64, 48
N_PARAMS = H * K + 2 * H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[:H * K].reshape(H, K)
    b1 = p[H * K:H * K + H]
    W2 = p[H * K + H:H * K + 2 * H]
    b2 = p[H * K + 2 * H]
    return W1, b1, W2, b2


def bf16_round(x):
    """Round-to-nearest-even to bfloat16 (as the device publishes W1)."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def logits(p, X, n_env, bf16_w1=False):
    W1, b1, W2, b2 = unpack(p)
    W1 = W1.copy()
    W1[:, n_env:] = 0.0
    if bf16_w1:
        W1 = bf16_round(W1).astype(np.float64)
    X = np.asarray(X, dtype=np.float64)
    h = X @ W1.T + b1
    hn = -X @ W1.T + b1
    lc = np.maximum(h, 0) @ W2 + b2
    ln = np.maximum(hn, 0) @ W2 + b2
    return np.stack([lc, ln], axis=1)


def loss_and_grad(p, X, y, n_env, t=None):
    W1, b1, W2, b2 = unpack(p)
    W1 = W1.copy()
    W1[:, n_env:] = 0.0
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    t = np.ones(len(y)) if t is None else np.asarray(t, dtype=np.float64)
    h = X @ W1.T + b1
    r = np.maximum(h, 0)
    z = (r @ W2 + b2) * t
    pz = 1.0 / (1.0 + np.exp(-z))
    # log sigmoid(z) = -log(1 + e^-z), log(1 - sigmoid(z)) = -log(1 + e^z)
    loss = np.mean(y * np.logaddexp(0.0, -z) + (1 - y) * np.logaddexp(0.0, z))
    n = len(y)
    dz = t * (pz - y) / n
    gW2 = r.T @ dz
    gb2 = dz.sum()
    dh = np.outer(dz, W2) * (h > 0)
    gW1 = dh.T @ X
    gW1[:, n_env:] = 0.0
    gb1 = dh.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2, [gb2]])
    return loss, g


def finite_difference(p, X, y, n_env, idx, eps=1e-6, t=None):
    g = []
    for i in idx:
        pp = np.array(p, dtype=np.float64)
        pm = pp.copy()
        pp[i] += eps
        pm[i] -= eps
        g.append((loss_and_grad(pp, X, y, n_env, t)[0] - loss_and_grad(pm, X, y, n_env, t)[0]) / (2 * eps))
    return np.array(g)

"""numpy oracle of the learned WCNS sensor (TEST INFRASTRUCTURE ONLY).

No counterpart in the reference.  Per interface a+1/2 of a grid line the
sensor reads 12 features of density and pressure at the six nodes
a-2 .. a+3 of the WCNS stencil and predicts whether the WENO5 nonlinear
weights differ from the linear ones (1/16, 10/16, 5/16) by more than
LABEL_BOUND for any of the ten interpolations of the face (5 conservative
variables, left and right states).

Features, per quantity phi in (rho, p), with v0..v5 = phi(a-2..a+3):
  beta_k = 13/12 (v_k - 2 v_k+1 + v_k+2)^2 + 1/4 (v_k - v_k+2)^2, k = 0..3
  S = sum_k beta_k
  f = beta_k / (S + 1e-30) (4), |v3 - v2| / (|v2| + |v3| + 1e-30),
      (log10(S + 1e-30) + 12) / 10
Network: 12 -> 32 (ReLU) -> 32 (ReLU) -> 1 (sigmoid), binary cross-entropy.
Parameter vector layout (1505 floats): W1[12][32], b1[32], W2[32][32],
b2[32], w3[32], b3.
"""
import numpy as np

NF = 12
H = 32
NPARAM = NF * H + H + H * H + H + H + 1
LABEL_BOUND = 1.0e-3
DLIN = np.array([1.0 / 16.0, 10.0 / 16.0, 5.0 / 16.0])


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    W1 = w[o:o + NF * H].reshape(NF, H); o += NF * H
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    w3 = w[o:o + H]; o += H
    b3 = w[o]
    return W1, b1, W2, b2, w3, b3


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((NF, H)) * np.sqrt(2.0 / NF)
    W2 = rng.standard_normal((H, H)) * np.sqrt(2.0 / H)
    w3 = rng.standard_normal(H) * np.sqrt(1.0 / H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]]).astype(np.float32)


def features(v_rho, v_p):
    """v_rho, v_p: arrays (..., 6) of node values a-2..a+3 -> (..., 12)."""
    out = []
    for v in (np.asarray(v_rho, dtype=np.float64), np.asarray(v_p, dtype=np.float64)):
        b = np.stack([13.0 / 12.0 * (v[..., k] - 2 * v[..., k + 1] + v[..., k + 2]) ** 2 +
                      0.25 * (v[..., k] - v[..., k + 2]) ** 2 for k in range(4)], axis=-1)
        S = b.sum(-1)
        out.append(b / (S[..., None] + 1e-30))
        out.append((np.abs(v[..., 3] - v[..., 2]) / (np.abs(v[..., 2]) + np.abs(v[..., 3]) + 1e-30))[..., None])
        out.append(((np.log10(S + 1e-30) + 12.0) / 10.0)[..., None])
    return np.concatenate(out, axis=-1)


def weno5_weights(a, b, c, d, e):
    """Nonlinear WCNS weights of the left-biased interpolation from
    u(a-2..a+2) (Jiang-Shu indicators, eps 1e-6)."""
    k = 13.0 / 12.0
    b0 = k * (a - 2 * b + c) ** 2 + 0.25 * (a - 4 * b + 3 * c) ** 2
    b1 = k * (b - 2 * c + d) ** 2 + 0.25 * (b - d) ** 2
    b2 = k * (c - 2 * d + e) ** 2 + 0.25 * (3 * c - 4 * d + e) ** 2
    al = np.stack([DLIN[0] / (1e-6 + b0) ** 2, DLIN[1] / (1e-6 + b1) ** 2, DLIN[2] / (1e-6 + b2) ** 2], -1)
    return al / al.sum(-1, keepdims=True)


def labels(qwin):
    """qwin: (..., 6, 5) conservative variables at a-2..a+3 -> 0/1 label:
    max over the ten interpolations of max_k |w_k - d_k| > LABEL_BOUND."""
    q = np.asarray(qwin, dtype=np.float64)
    dev = np.zeros(q.shape[:-2])
    for x in range(5):
        v = q[..., x]
        wl = weno5_weights(v[..., 0], v[..., 1], v[..., 2], v[..., 3], v[..., 4])
        wr = weno5_weights(v[..., 5], v[..., 4], v[..., 3], v[..., 2], v[..., 1])
        dev = np.maximum(dev, np.abs(wl - DLIN).max(-1))
        dev = np.maximum(dev, np.abs(wr - DLIN).max(-1))
    return (dev > LABEL_BOUND).astype(np.float32)


def forward(w, X):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    z1 = X @ W1 + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2 + b2
    a2 = np.maximum(z2, 0.0)
    z3 = a2 @ w3 + b3
    return z1, a1, z2, a2, z3


def loss_and_grad(w, X, y):
    """Mean binary cross-entropy of sigmoid(z3) and its gradient (float64)."""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(w)
    z1, a1, z2, a2, z3 = forward(w, X)
    n = X.shape[0]
    # BCE with logits: softplus(z) - y z
    loss = np.mean(np.logaddexp(0.0, z3) - y * z3)
    p = 1.0 / (1.0 + np.exp(-z3))
    d3 = (p - y) / n
    gw3 = a2.T @ d3
    gb3 = d3.sum()
    d2 = np.outer(d3, w3) * (z2 > 0)
    gW2 = a1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def predict(w, X):
    return 1.0 / (1.0 + np.exp(-forward(w, X)[-1]))

"""NumPy reference of the step-proposer MLP (9 -> 64 -> 64 -> 64 -> 2, ReLU).

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward / backward kernels.
Parameter layout (one flat FP32 vector): W1[9][64] b1[64] W2[64][64] b2[64]
W3[64][64] b3[64] W4[64][2] b4[2].  Loss: mean_b (w0 (y0-t0)^2 + w1 (y1-t1)^2).
"""
import numpy as np

IN, H, OUT = 9, 64, 2
SHAPES = [(IN, H), (H,), (H, H), (H,), (H, H), (H,), (H, OUT), (OUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def init(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for s in SHAPES:
        if len(s) == 2:
            parts.append(rng.normal(0.0, np.sqrt(2.0 / s[0]), size=s).ravel())
        else:
            parts.append(np.zeros(s))
    return np.concatenate(parts).astype(np.float32)


def forward(p, X):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    z1 = X @ W1 + b1; y1 = np.maximum(z1, 0)
    z2 = y1 @ W2 + b2; y2 = np.maximum(z2, 0)
    z3 = y2 @ W3 + b3; y3 = np.maximum(z3, 0)
    y4 = y3 @ W4 + b4
    return y4, (z1, y1, z2, y2, z3, y3)


def loss_grad(p, X, T, w0=1.0, w1=1.0):
    """loss and analytic gradient (flat, same layout as p), float64"""
    X = np.asarray(X, np.float64); T = np.asarray(T, np.float64)
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y4, (z1, y1, z2, y2, z3, y3) = forward(p, X)
    B = X.shape[0]
    w = np.array([w0, w1])
    e = y4 - T
    loss = float(np.sum(w * e * e) / B)
    d4 = 2.0 * w * e / B
    gW4 = y3.T @ d4; gb4 = d4.sum(0)
    d3 = (d4 @ W4.T) * (z3 > 0)
    gW3 = y2.T @ d3; gb3 = d3.sum(0)
    d2 = (d3 @ W3.T) * (z2 > 0)
    gW2 = y1.T @ d2; gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1 = X.T @ d1; gb1 = d1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g


def loss_only(p, X, T, w0=1.0, w1=1.0):
    y4, _ = forward(p, np.asarray(X, np.float64))
    e = y4 - np.asarray(T, np.float64)
    return float(np.sum(np.array([w0, w1]) * e * e) / X.shape[0])

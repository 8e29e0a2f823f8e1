"""ORACLE — test infrastructure only.

A numpy restatement of the order-cost MLP's forward and backward passes
(gem_amd/csrc/hip/mlp.hip), used to check the GPU gradients.  Same math: input
x = features / 64, ReLU layers with bf16-rounded weights (round to nearest
even), fp32 biases, loss = mean((y - t)^2) / 2; gradients are taken with
respect to the bf16 weights actually used.  Accumulation is in float64 here,
so GPU/CPU differences are bounded by fp32 accumulation error only.
"""
import numpy as np

FEAT_SCALE = 1.0 / 64.0


def bf16_round(a):
    """Round float32 values to bfloat16 precision (nearest even), as float32."""
    x = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((x & 0x7F800000) == 0x7F800000) & ((x & 0x007FFFFF) != 0)
    r = (x + 0x7FFF + ((x >> 16) & 1)) >> 16
    r = np.where(nan, (x >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def unpack(p, in_dim, hidden):
    o = 0
    W1 = p[o:o + in_dim * hidden].reshape(in_dim, hidden); o += in_dim * hidden
    b1 = p[o:o + hidden]; o += hidden
    W2 = p[o:o + hidden * hidden].reshape(hidden, hidden); o += hidden * hidden
    b2 = p[o:o + hidden]; o += hidden
    W3 = p[o:o + hidden]; o += hidden
    b3 = p[o:o + 1]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, feats, in_dim, hidden, round_weights=True):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64 if not round_weights else np.float32), in_dim, hidden)
    if round_weights:
        W1b, W2b, W3b = (bf16_round(w).astype(np.float64) for w in (W1, W2, W3))
        b1, b2, b3 = (np.asarray(b, np.float64) for b in (b1, b2, b3))
    else:
        W1b, W2b, W3b = W1, W2, W3
    X = np.asarray(feats, np.float64) * FEAT_SCALE
    H1 = np.maximum(X @ W1b + b1, 0.0)
    H2 = np.maximum(H1 @ W2b + b2, 0.0)
    Y = H2 @ W3b + b3[0]
    return X, H1, H2, Y, (W1b, W2b, W3b)


def loss_grad(p, feats, targets, in_dim, hidden, round_weights=True):
    X, H1, H2, Y, (W1b, W2b, W3b) = forward(p, feats, in_dim, hidden, round_weights)
    n = len(Y)
    d = Y - np.asarray(targets, np.float64)
    loss = 0.5 * np.sum(d * d) / n
    dY = d / n
    gW3 = H2.T @ dY
    gb3 = np.array([dY.sum()])
    dH2 = np.outer(dY, W3b) * (H2 > 0)
    gW2 = H1.T @ dH2
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ W2b.T) * (H1 > 0)
    gW1 = X.T @ dH1
    gb1 = dH1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
    return loss, g

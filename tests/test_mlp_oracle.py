"""Pins for the MLP oracle (oracle/mlp.py): bf16 rounding against torch's
bfloat16 conversion, and the backward pass against finite differences and
torch autograd in float64."""
import numpy as np
import torch

from oracle import mlp


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10 ** rng.uniform(-6, 6, 10000),
                        np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -2.5, 0.0, 65504.0], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mlp.bf16_round(x), want)


def _tiny(seed=1, in_dim=12, hidden=8, n=20):
    rng = np.random.default_rng(seed)
    P = in_dim * hidden + hidden + hidden * hidden + hidden + hidden + 1
    p = rng.standard_normal(P) * 0.3
    feats = rng.integers(0, 200, size=(n, in_dim))
    t = rng.standard_normal(n)
    return p, feats, t, in_dim, hidden


def test_backward_vs_finite_differences():
    p, feats, t, F, H = _tiny()
    loss, g = mlp.loss_grad(p, feats, t, F, H, round_weights=False)
    eps = 1e-6
    rng = np.random.default_rng(2)
    for i in rng.choice(p.size, 60, replace=False):
        q = p.copy()
        q[i] += eps
        lp, _ = mlp.loss_grad(q, feats, t, F, H, round_weights=False)
        q[i] -= 2 * eps
        lm, _ = mlp.loss_grad(q, feats, t, F, H, round_weights=False)
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-6 * max(1.0, abs(g[i]))


def test_backward_vs_torch_autograd():
    p, feats, t, F, H = _tiny(seed=3)
    loss, g = mlp.loss_grad(p, feats, t, F, H, round_weights=False)
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    o = 0
    W1 = P[o:o + F * H].view(F, H); o += F * H
    b1 = P[o:o + H]; o += H
    W2 = P[o:o + H * H].view(H, H); o += H * H
    b2 = P[o:o + H]; o += H
    W3 = P[o:o + H]; o += H
    b3 = P[o]
    X = torch.tensor(feats, dtype=torch.float64) * mlp.FEAT_SCALE
    Y = torch.relu(torch.relu(X @ W1 + b1) @ W2 + b2) @ W3 + b3
    L = 0.5 * ((Y - torch.tensor(t)) ** 2).mean()
    L.backward()
    assert abs(L.item() - loss) < 1e-12
    assert np.allclose(P.grad.numpy(), g, rtol=1e-10, atol=1e-12)

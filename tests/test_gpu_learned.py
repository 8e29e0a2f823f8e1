"""GPU tests of the learned join order.

The MLP's forward pass and gradients are checked against the numpy oracle;
training must reduce the loss; model-chosen orders must leave every count
unchanged; and on held-out queries the trained model's orders are compared
with GEM's (the gain is printed and must not be a loss)."""
import numpy as np
import pytest

from oracle import mlp as mlp_oracle
from gem_amd import Graph, Query
from gem_amd.learned import OrderModel, choose, collect, match_batch_learned

pytestmark = pytest.mark.gpu

F = 272
H = 128


def test_mlp_forward_and_gradients_match_oracle(gpu_ctx):
    model = OrderModel(gpu_ctx, seed=3)
    rng = np.random.default_rng(0)
    feats = rng.integers(0, 600, size=(300, F)).astype(np.int32)
    targets = rng.uniform(0, 30, 300).astype(np.float32)
    p = model.get_params()
    pred = model.predict(feats)
    _, _, _, Y, _ = mlp_oracle.forward(p, feats, F, H)
    assert np.allclose(pred, Y, rtol=1e-4, atol=1e-4 * np.abs(Y).max())
    loss, g = model.loss_grad(feats, targets)
    rloss, rg = mlp_oracle.loss_grad(p, feats, targets, F, H)
    assert abs(loss - rloss) <= 1e-4 * abs(rloss)
    scale = np.abs(rg).max()
    assert np.max(np.abs(g - rg)) <= 1e-4 * scale, np.max(np.abs(g - rg)) / scale


def test_training_reduces_loss(gpu_ctx):
    model = OrderModel(gpu_ctx, seed=5)
    rng = np.random.default_rng(1)
    feats = rng.integers(0, 600, size=(512, F)).astype(np.int32)
    w = rng.standard_normal(F) / F
    targets = (feats.astype(np.float64) / 64 @ w + 10).astype(np.float32)
    losses = model.train(feats, targets, iters=200, lr=1e-3)
    assert losses[-1] < 0.05 * losses[0]


@pytest.fixture(scope="module")
def rmat_graph():
    return Graph.rmat(14, 16.0, 8, 1002)


def _queries(g, n, ks, seed):
    out, i = [], 0
    while len(out) < n:
        q = Query.random_walk(g, ks[i % len(ks)], seed + i)
        i += 1
        if q.num_e >= 3:
            out.append(q)
    return out


def test_learned_orders_gain_and_keep_counts(gpu_ctx, rmat_graph):
    gpu_ctx.load_data(rmat_graph)
    cap, wl = 10 ** 8, 3 * 10 ** 8
    train_q = _queries(rmat_graph, 48, [5, 6, 7, 8], 20001)
    test_q = _queries(rmat_graph, 24, [5, 6, 7, 8], 30001)
    Ftr, Ytr, Qtr, Ctr, _ = collect(gpu_ctx, train_q, max_orders=12, cap=cap, work_limit=wl)
    model = OrderModel(gpu_ctx, seed=11)
    mu, sd = float(Ytr.mean()), float(Ytr.std() + 1e-6)
    losses = model.train(Ftr, (Ytr - mu) / sd, iters=400, lr=2e-3)
    assert losses[-1] < losses[0]
    Fte, Yte, Qte, Cte, CNTte = collect(gpu_ctx, test_q, max_orders=12, cap=cap, work_limit=wl)
    pick = choose(model, Fte, Qte)
    gem_cost = sum(2.0 ** Yte[(Qte == q) & (Cte == 0)][0] for q in range(len(test_q)))
    learned_cost = sum(2.0 ** Yte[pick[q]] for q in range(len(test_q)))
    best_cost = sum(2.0 ** Yte[Qte == q].min() for q in range(len(test_q)))
    print(f"held-out rows processed: GEM {gem_cost:.3e}  learned {learned_cost:.3e}  best candidate {best_cost:.3e}"
          f"  gain {gem_cost / learned_cost:.2f}x", flush=True)
    assert learned_cost <= gem_cost * 1.05
    # every finished candidate order of a query gives the same count
    for q in range(len(test_q)):
        c = CNTte[(Qte == q) & (CNTte >= 0)]
        assert len(set(c.tolist())) <= 1
    # the batched path with the model: counts equal to GEM order's
    gem = match_batch_learned(gpu_ctx, test_q, None, cap=cap)
    lrn = match_batch_learned(gpu_ctx, test_q, model, max_orders=12, cap=cap)
    for a, b in zip(gem, lrn):
        if not a["work_limited"] and not b["work_limited"]:
            assert (a["count"], a["saturated"]) == (b["count"], b["saturated"])


def test_model_outliving_its_context():
    """Destroying a context before its model releases the model's device
    memory; the model then reports GEM_E_STATE and can still be destroyed."""
    import ctypes as C
    from gem_amd import GemContext
    from gem_amd._lib import load
    lib = load()
    ctx = GemContext(0)
    h = C.c_void_p()
    assert lib.gem_mlp_create(ctx._h, F, H, 1, C.byref(h)) == 0
    lib.gem_hip_destroy(ctx._h)   # context first, through the C ABI
    ctx._h = None
    feats = np.zeros((2, F), np.int32)
    out = np.zeros(2, np.float32)
    st = lib.gem_mlp_predict(h, feats.ctypes.data_as(C.POINTER(C.c_int32)), 2,
                             out.ctypes.data_as(C.POINTER(C.c_float)))
    assert st == 5  # GEM_E_STATE
    lib.gem_mlp_destroy(h)
    # the Python wrappers close models before their context
    ctx2 = GemContext(0)
    m = OrderModel(ctx2, seed=2)
    ctx2.close()
    assert m._h is None

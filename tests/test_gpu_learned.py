"""GPU tests of the learned join order.

The MLP's forward pass and gradients are checked against the numpy oracle;
training must reduce the loss; model-ranked orders must leave every count
unchanged; and on queries held out from training, the restart schedule must
do no more work with the model's ranking than with the candidates'
generation order (GEM's first): the shipped R-MAT 2^20 model on the
learned-order workload's 256 queries (its training queries are another
seed's), and a model trained here from scratch on a small graph."""
import numpy as np
import pytest

from oracle import mlp as mlp_oracle
from gem_amd import Graph, Query
import os

from gem_amd.learned import OrderModel, collect, match_batch_learned, rank_orders, relative_targets, schedule_work
from gem_amd.workloads import make_graph, make_queries

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


def _schedule(T, ranking, first, limit):
    tot = 0.0
    for q in np.unique(T["query"]):
        m = T["query"] == q
        costs = dict(zip(T["cand"][m].tolist(), np.where(T["limited"][m], np.inf, T["work"][m]).tolist()))
        tot += schedule_work(costs, ranking[int(q)], first, 4, 1, limit)
    return tot


def test_learned_orders_gain_and_keep_counts(gpu_ctx, rmat_graph):
    """Trained here on 48 queries of a 2^14-vertex R-MAT graph, evaluated on 24
    others: the schedule's work with the model's ranking is printed, and the
    batched path's counts with the model equal those without it."""
    gpu_ctx.load_data(rmat_graph)
    cap, wl, first = 10 ** 8, 3 * 10 ** 8, 10 ** 4
    train_q = _queries(rmat_graph, 48, [5, 6, 7, 8], 20001)
    test_q = _queries(rmat_graph, 24, [5, 6, 7, 8], 30001)
    Ttr = collect(gpu_ctx, train_q, max_orders=8, cap=cap, work_limit=wl)
    Y = relative_targets(Ttr["work"], Ttr["query"])
    model = OrderModel(gpu_ctx, seed=11)
    mu, sd = float(Y.mean()), float(Y.std() + 1e-6)
    losses = model.train(Ttr["feats"], (Y - mu) / sd, iters=400, lr=2e-3)
    assert losses[-1] < losses[0]
    Tte = collect(gpu_ctx, test_q, max_orders=8, cap=cap, work_limit=wl)
    pred = model.predict(Tte["feats"]) * sd + mu
    learned = _schedule(Tte, rank_orders(pred, Tte["query"], Tte["cand"]), first, wl)
    gen = _schedule(Tte, {int(q): sorted(Tte["cand"][Tte["query"] == q].tolist()) for q in np.unique(Tte["query"])},
                    first, wl)
    print(f"held-out schedule work: generation order {gen:.3e}  learned {learned:.3e}  gain {gen / learned:.2f}x",
          flush=True)
    # every finished candidate order of a query gives the same count
    for q in np.unique(Tte["query"]):
        c = Tte["count"][(Tte["query"] == q) & (Tte["count"] >= 0)]
        assert len(set(c.tolist())) <= 1
    # the batched path with the model: counts equal to those without it
    base = match_batch_learned(gpu_ctx, test_q, None, cap=cap)
    lrn = match_batch_learned(gpu_ctx, test_q, model, max_orders=8, cap=cap, first_attempt=first)
    for a, b in zip(base, lrn):
        assert not b["work_limited"]
        if not a["work_limited"]:
            assert (a["count"], a["saturated"]) == (b["count"], b["saturated"])


def test_shipped_model_gains_on_heldout_workload(gpu_ctx):
    """The shipped R-MAT 2^20 order model (trained on rmat20_train queries)
    on the 256 learned-order queries of 8-24 vertices: with restarts, every
    query ends, counts equal the run without the model, and the model's
    ranking does no more work than the generation order (gain >= 1)."""
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gem_amd", "models",
                        "order_mlp_rmat20.npz")
    g = make_graph("rmat20_learn")
    qs = make_queries("rmat20_learn", g)
    gpu_ctx.load_data(g)
    model = OrderModel.load(gpu_ctx, path)
    assert model.meta["workload"] == "rmat20_train"
    kw = dict(max_orders=8, cap=10 ** 9, first_attempt=10 ** 8)
    gpu_ctx.set_racing(False)  # the ranking's own effect: no twins racing stragglers
    try:
        base = match_batch_learned(gpu_ctx, qs, None, **kw)
        lrn = match_batch_learned(gpu_ctx, qs, model, **kw)
    finally:
        gpu_ctx.set_racing(True)
    assert not any(r["work_limited"] for r in base + lrn)
    assert [(r["count"], r["saturated"]) for r in base] == [(r["count"], r["saturated"]) for r in lrn]
    wb, wl = sum(r["work"] for r in base), sum(r["work"] for r in lrn)
    print(f"rmat20_learn: work without the model {wb:.3e}, with it {wl:.3e}, gain {wb / wl:.2f}x", flush=True)
    assert wb / wl >= 1.0
    model.close()


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

"""Learned join order: the order-cost MLP (GPU kernels in libgem_hip.so),
its training data and the restart schedule it ranks orders for.

Training data are tuples the matcher logs itself: for each training query the
library proposes candidate vertex orders (gem_hip_order_candidates: GEM's
order first, then estimate-greedy orders, each also with the query's leaves
moved last, then seeded random orders), the batched vertex-at-a-time matcher
runs every (query, order) pair at the benchmark's count cap, and logs the
rows each level produced (the intermediate-result sizes) and the work it did
(rows read plus list entries walked, the cost the restart limits bound).
The label of a tuple is log2(1 + work) minus the least such value over its
query's orders, so the network learns how the orders of one query compare.
The network is fit by full-batch Adam on the GPU with the library's own
forward/backward kernels.  At match time the batched path ranks each query's
candidates by predicted cost and runs them in that order under geometrically
growing work limits (restarts), so a poor first choice costs a bounded amount
of work; without a model the candidates run in generation order (GEM's
first).
"""
import ctypes as C

import numpy as np

from ._lib import GEM_FEAT_DIM, GEM_MAX_QV, BatchOpts, GemError, JoinResult, QueryView, load

HIDDEN = 128


def _pi32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _pf32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class OrderModel:
    """3-layer MLP (GEM_FEAT_DIM -> 128 -> 128 -> 1) living on the context's GPU."""

    def __init__(self, ctx, seed=0, in_dim=GEM_FEAT_DIM, hidden=HIDDEN):
        self._lib = load()
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(self._lib.gem_mlp_create(ctx._h, in_dim, hidden, seed, C.byref(h)), "gem_mlp_create")
        self._h = h
        self.in_dim, self.hidden = in_dim, hidden
        self.num_params = int(self._lib.gem_mlp_num_params(h))
        ctx._models.add(self)  # closed before the context is

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.gem_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @classmethod
    def load(cls, ctx, path):
        """A model saved by gem_amd.train_order_model."""
        z = np.load(path, allow_pickle=False)
        m = cls(ctx, seed=0, in_dim=int(z["in_dim"]), hidden=int(z["hidden"]))
        m.set_params(z["params"])
        m.meta = {k: z[k].item() if z[k].shape == () else z[k] for k in z.files if k != "params"}
        m.set_margin(float(m.meta.get("margin", 0.0)))
        return m

    def set_margin(self, log2_units):
        """GEM's order is replaced only by a candidate predicted cheaper by
        more than `log2_units` (log2 of the work; 1.0 = half the work)."""
        sd = float(getattr(self, "meta", {}).get("sd", 1.0))
        self.margin = float(log2_units)
        self.ctx._check(self._lib.gem_mlp_set_margin(self._h, self.margin / sd), "gem_mlp_set_margin")

    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        self.ctx._check(self._lib.gem_mlp_get_params(self._h, _pf32(out)), "gem_mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        self.ctx._check(self._lib.gem_mlp_set_params(self._h, _pf32(p)), "gem_mlp_set_params")

    def predict(self, feats):
        f = np.ascontiguousarray(feats, dtype=np.int32).reshape(-1, self.in_dim)
        out = np.zeros(len(f), np.float32)
        if len(f):
            self.ctx._check(self._lib.gem_mlp_predict(self._h, _pi32(f), len(f), _pf32(out)), "gem_mlp_predict")
        return out

    def loss_grad(self, feats, targets):
        f = np.ascontiguousarray(feats, dtype=np.int32).reshape(-1, self.in_dim)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        loss = C.c_float()
        g = np.zeros(self.num_params, np.float32)
        self.ctx._check(self._lib.gem_mlp_loss_grad(self._h, _pi32(f), _pf32(t), len(f), C.byref(loss), _pf32(g)),
                        "gem_mlp_loss_grad")
        return float(loss.value), g

    def train(self, feats, targets, iters=300, lr=1e-3):
        f = np.ascontiguousarray(feats, dtype=np.int32).reshape(-1, self.in_dim)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        losses = np.zeros(iters, np.float32)
        self.ctx._check(self._lib.gem_mlp_train(self._h, _pi32(f), _pf32(t), len(f), iters, lr, _pf32(losses)),
                        "gem_mlp_train")
        return losses


def order_candidates(ctx, query, max_orders=16, seed=0):
    """Candidate vertex orders and their features for the context's current
    query (after filter + build_edge_tables)."""
    lib = load()
    orders = np.zeros(max_orders * query.num_v, np.uint32)
    feats = np.zeros(max_orders * GEM_FEAT_DIM, np.int32)
    n = C.c_uint32()
    ctx._check(lib.gem_hip_order_candidates(ctx._h, max_orders, seed, orders.ctypes.data_as(C.POINTER(C.c_uint32)),
                                            _pi32(feats), C.byref(n)), "gem_hip_order_candidates")
    k = n.value
    return orders[:k * query.num_v].reshape(k, query.num_v), feats[:k * GEM_FEAT_DIM].reshape(k, GEM_FEAT_DIM)


def cost_label(work):
    """log2(1 + work), work = rows read plus list entries walked by the
    vertex-at-a-time join."""
    return float(np.log2(1.0 + float(work)))


def collect(ctx, queries, max_orders=8, cap=0, work_limit=0, group=64, log=None):
    """Run every candidate order of every query through the batched matcher
    (group (query, order) pairs per run) and log the tuples.  Returns a dict
    of arrays over tuples: feats [N x D] int32, work [N] (rows read plus list
    entries walked; work_limit when stopped there), limited [N] bool, level_rows
    [N x GEM_MAX_QV] (rows after each level), count [N] (-1 when stopped),
    query [N], cand [N] (candidate index; 0 = GEM's order), nv [N]."""
    pairs, F, Q, Ci = [], [], [], []
    for qi, q in enumerate(queries):
        ctx.filter(q, want_lists=False)
        ctx.build_edge_tables()
        orders, feats = order_candidates(ctx, q, max_orders, 0)
        for ci in range(len(orders)):
            pairs.append((q, orders[ci]))
            F.append(feats[ci])
            Q.append(qi)
            Ci.append(ci)
    W, LIM, LR, CNT, NV = [], [], [], [], []
    for b in range(0, len(pairs), group):
        part = pairs[b:b + group]
        res = ctx.match_batch([p[0] for p in part], vertex_orders=[p[1] for p in part], cap=cap,
                              work_limit=work_limit, group=group)
        for (q, _), r in zip(part, res):
            W.append(min(r["work"], work_limit) if (work_limit and r["work_limited"]) else r["work"])
            LIM.append(bool(r["work_limited"]))
            lr = np.zeros(GEM_MAX_QV, np.uint64)
            lr[:len(r["level_rows"])] = r["level_rows"]
            LR.append(lr)
            CNT.append(-1 if r["work_limited"] else r["count"])
            NV.append(q.num_v)
        if log:
            log(f"[learn] ran {min(b + group, len(pairs))}/{len(pairs)} (query, order) pairs")
    return {"feats": np.array(F, np.int32).reshape(-1, GEM_FEAT_DIM), "work": np.array(W, np.float64),
            "limited": np.array(LIM, bool), "level_rows": np.array(LR, np.uint64).reshape(-1, GEM_MAX_QV),
            "count": np.array(CNT, np.int64), "query": np.array(Q, np.int64), "cand": np.array(Ci, np.int64),
            "nv": np.array(NV, np.int64)}


def relative_targets(work, query):
    """Per tuple: log2(1 + work) minus the least value over its query's tuples."""
    y = np.log2(1.0 + np.asarray(work, np.float64))
    best = {}
    for q, v in zip(query, y):
        best[q] = min(best.get(q, np.inf), v)
    return (y - np.array([best[q] for q in query])).astype(np.float32)


def rank_orders(pred, query, cand, margin=0.0):
    """Per query: candidate indices by predicted cost (ascending, ties by
    index); GEM's order (candidate 0) stays first unless the cheapest
    prediction beats it by more than `margin` (model output units).  The
    rule gem_hip_match_batch_learned applies."""
    out = {}
    for q in np.unique(query):
        m = np.flatnonzero(query == q)
        c, p = cand[m], pred[m]
        r = [int(c[i]) for i in np.argsort(p, kind="stable")]
        p0 = p[c == 0][0]
        if not p.min() < p0 - margin:
            r.remove(0)
            r.insert(0, 0)
        out[int(q)] = r
    return out


def schedule_work(costs, ranking, first_attempt, growth=4, width=1, limit=None):
    """Work the restart schedule spends on one query: attempt a runs
    ranking[a % n] and may do first_attempt * growth^(a // width) work; the
    first attempt whose order costs at most that ends the query.  costs[c] =
    work of candidate c run to its end (>= limit when the collection stopped
    it: such an order counts as never ending)."""
    spent = 0.0
    n = len(ranking)
    for a in range(64):
        c = ranking[a % n]
        budget = float(first_attempt) * float(growth) ** (a // max(width, 1))
        w = costs[c]
        if (limit is None or w < limit) and w <= budget:
            return spent + w
        spent += budget
    return spent


def match_batch_learned(ctx, queries, model=None, max_orders=16, cap=0, work_limit=0, level_budget=0, group=0,
                        first_attempt=0, restart_growth=4, restart_width=1, max_attempts=0):
    """Batched path with model-ranked orders (model None = candidates in
    generation order, GEM's first).  first_attempt > 0: attempt a of a query
    runs its (a mod n)-th ranked order and may do first_attempt *
    restart_growth^(a // restart_width) work before the query starts again
    (see gem_hip_match_batch_learned); work_limit bounds the total."""
    lib = load()
    nq = len(queries)
    views = (QueryView * max(nq, 1))(*[q.view() for q in queries])
    opts = BatchOpts(int(cap), int(work_limit), int(level_budget), int(group), int(restart_growth), int(first_attempt),
                     int(restart_width), int(max_attempts))
    res = (JoinResult * max(nq, 1))()
    chosen = np.zeros(max(nq, 1), np.uint32)
    ctx._check(lib.gem_hip_match_batch_learned(ctx._h, views, nq, model._h if model is not None else None,
                                               max_orders, C.byref(opts), res, None,
                                               chosen.ctypes.data_as(C.POINTER(C.c_uint32))),
               "gem_hip_match_batch_learned")
    return [{"count": int(res[i].count), "saturated": bool(res[i].saturated),
             "work_limited": bool(res[i].work_limited), "rows_processed": int(res[i].rows_processed),
             "work": int(res[i].work), "restarts": int(res[i].restarts), "chosen": int(chosen[i])}
            for i in range(nq)]


__all__ = ["OrderModel", "order_candidates", "collect", "relative_targets", "rank_orders", "schedule_work",
           "match_batch_learned", "cost_label", "GemError"]

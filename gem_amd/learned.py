"""Learned join order: the order-cost MLP (GPU kernels in libgem_hip.so) and
its training loop.

Training data are (order features, cost) pairs the matcher logs itself: for
each training query the library proposes candidate vertex orders
(gem_hip_order_candidates: GEM's order first, then estimate-greedy and seeded
random orders), the vertex-at-a-time matcher runs each one, and the cost
label is log2(1 + work), work = rows read plus list entries walked.  The
network is fit by full-batch Adam on the GPU; at match time the batched path
runs the candidate with the lowest predicted cost when it beats GEM's order by
more than the model's margin, else GEM's order (also the order when no model
is given).
"""
import ctypes as C

import numpy as np

from ._lib import GEM_FEAT_DIM, BatchOpts, GemError, JoinResult, QueryView, load

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
    """Training label of one (query, order) run: log2(1 + work), work = rows
    read plus list entries walked by the vertex-at-a-time join."""
    return float(np.log2(1.0 + float(work)))


def collect(ctx, queries, max_orders=16, cap=0, work_limit=0, seed=0, log=None):
    """Run every candidate order of every query with the vertex-at-a-time
    matcher and log (features, cost).  Returns (feats [N x D] int32,
    costs [N] float32, query index [N], candidate index [N], counts [N])."""
    F, Y, QI, CI, CNT = [], [], [], [], []
    old_limit = None
    if work_limit:
        ctx.set_work_limit(work_limit)
    for qi, q in enumerate(queries):
        ctx.filter(q, want_lists=False)
        ctx.build_edge_tables()
        orders, feats = order_candidates(ctx, q, max_orders, seed + qi)
        for ci in range(len(orders)):
            r = ctx.vjoin(orders[ci], cap=cap)
            F.append(feats[ci])
            Y.append(cost_label(r["work"]))
            QI.append(qi)
            CI.append(ci)
            CNT.append(-1 if r["work_limited"] else r["count"])
        if log and (qi + 1) % 16 == 0:
            log(f"[learn] logged {qi + 1}/{len(queries)} queries, {len(Y)} tuples")
    if work_limit:
        ctx.set_work_limit(0)
    del old_limit
    return (np.array(F, np.int32).reshape(-1, GEM_FEAT_DIM), np.array(Y, np.float32), np.array(QI),
            np.array(CI), np.array(CNT, np.int64))


def choose(model, feats, qidx, margin=0.0):
    """Per query, the candidate with the lowest predicted cost, if it beats the
    query's first candidate (GEM's order) by more than `margin` (model output
    units); else the first candidate."""
    pred = model.predict(feats)
    first, best = {}, {}
    for i, (q, p) in enumerate(zip(qidx, pred)):
        if q not in first:
            first[q] = (i, p)
        if q not in best or p < best[q][1]:
            best[q] = (i, p)
    return {q: (best[q][0] if best[q][1] < first[q][1] - margin else first[q][0]) for q in best}


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


__all__ = ["OrderModel", "order_candidates", "collect", "choose", "match_batch_learned", "cost_label", "GemError"]

"""Train the order-cost MLP on costs the matcher logs itself.

    python -m gem_amd.train_order_model --workload rmat20_learn --out gem_amd/models/order_mlp_rmat20.npz

For every training query the library proposes candidate vertex orders (GEM's
order first) with their int32 features; every (query, order) pair is run by
the batched vertex-at-a-time matcher (many pairs per launch), which logs the
rows each level produced and the work it did (rows read plus list entries
walked); the cost label is log2(1 + work).  The MLP is then fit on the GPU
(full-batch Adam, the library's own forward/backward kernels) on a random 70%
of the queries, stopped at the step count whose choices do the least work on
another 15% (checked every --eval-every steps), and evaluated on the last 15%:
total work of the model's choice vs GEM's order.  With --val-frac 0 every step
runs on 80% and the other 20% are the evaluation set (the shipped rmat22 model:
--val-frac 0 --iters 600 --lr 2e-3 on the 256 default queries).  The weights, the label
normalisation and the evaluation are saved to an .npz file.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import GemContext
from ._lib import GEM_FEAT_DIM
from .learned import OrderModel, cost_label, order_candidates
from .workloads import CAP, make_graph, make_queries


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def collect_batched(ctx, queries, max_orders, cap, work_limit, group, seed=0):
    """(features, costs, query index, candidate index, counts) for every
    candidate order of every query, run through the batched path."""
    entries, feats, qidx, cidx = [], [], [], []
    for qi, q in enumerate(queries):
        ctx.filter(q, want_lists=False)
        ctx.build_edge_tables()
        orders, f = order_candidates(ctx, q, max_orders, seed + qi)
        for ci in range(len(orders)):
            entries.append((q, orders[ci]))
            feats.append(f[ci])
            qidx.append(qi)
            cidx.append(ci)
    log(f"[train] {len(entries)} (query, order) pairs from {len(queries)} queries")
    costs, counts = [], []
    t = time.perf_counter()
    for b in range(0, len(entries), group):
        part = entries[b:b + group]
        res = ctx.match_batch([e[0] for e in part], vertex_orders=[e[1] for e in part], cap=cap,
                              work_limit=work_limit, group=group)
        for r in res:
            costs.append(cost_label(r["work"]))
            counts.append(-1 if r["work_limited"] else r["count"])
        log(f"[train]   ran {min(b + group, len(entries))}/{len(entries)} pairs, {time.perf_counter() - t:.1f}s")
    return (np.array(feats, np.int32).reshape(-1, GEM_FEAT_DIM), np.array(costs, np.float32), np.array(qidx),
            np.array(cidx), np.array(counts, np.int64))


def evaluate(model, F, Y, Q, C, mu, sd, queries_sel, margin=0.0):
    """Total work (2^label) over the selected queries of GEM's order, of the
    model's choice (GEM's order unless a candidate is predicted cheaper by more
    than `margin` log2 units) and of the best candidate."""
    pred = model.predict(F) * sd + mu
    gem = learned = best = 0.0
    for q in queries_sel:
        m = Q == q
        idx = np.flatnonzero(m)
        g0 = idx[C[idx] == 0][0]
        pick = idx[np.argmin(pred[idx])]
        if not pred[pick] < pred[g0] - margin:
            pick = g0
        gem += 2.0 ** Y[g0]
        learned += 2.0 ** Y[pick]
        best += 2.0 ** Y[idx].min()
    return {"gem_work": float(gem), "learned_work": float(learned), "best_work": float(best),
            "gain_vs_gem": float(gem / learned) if learned else 0.0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="rmat20_learn")
    ap.add_argument("--queries", type=int, default=0, help="use only the first N queries")
    ap.add_argument("--num-queries", type=int, default=0, help="queries generated for the workload (0 = its default)")
    ap.add_argument("--eval-every", type=int, default=50, help="Adam steps between validation checks")
    ap.add_argument("--val-frac", type=float, default=0.15,
                    help="share of the queries that picks the step count (0: train all --iters steps on 80%%, "
                         "evaluate on 20%%)")
    ap.add_argument("--orders", type=int, default=16)
    # the labels only have to rank a query's orders: a lower cap and work limit
    # than the benchmark's keep the collection cheap
    ap.add_argument("--cap", type=float, default=CAP / 100)
    ap.add_argument("--work-limit", type=float, default=3e8)
    ap.add_argument("--group", type=int, default=32)
    ap.add_argument("--cache", default=None, help=".npz of collected tuples: loaded if present, else written")
    ap.add_argument("--iters", type=int, default=1000, help="most Adam steps")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--margin", type=float, default=1.0,
                    help="log2 units by which a candidate must beat GEM's predicted cost to replace it")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "models", "order_mlp_rmat20.npz"))
    args = ap.parse_args()

    g = make_graph(args.workload)
    qs = make_queries(args.workload, g, args.num_queries or None)
    if args.queries:
        qs = qs[:args.queries]
    ctx = GemContext(0)
    ctx.load_data(g)
    if args.cache and os.path.exists(args.cache):
        z = np.load(args.cache, allow_pickle=False)
        F, Y, Q, C, CNT = z["F"], z["Y"], z["Q"], z["C"], z["CNT"]
        log(f"[train] loaded {len(Y)} tuples from {args.cache}")
    else:
        F, Y, Q, C, CNT = collect_batched(ctx, qs, args.orders, int(args.cap), int(args.work_limit), args.group)
        if args.cache:
            np.savez(args.cache, F=F, Y=Y, Q=Q, C=C, CNT=CNT)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(len(qs))
    nva = int(args.val_frac * len(qs))
    ntr = int(0.7 * len(qs)) if nva else int(0.8 * len(qs))
    tr_q, va_q, te_q = set(perm[:ntr].tolist()), perm[ntr:ntr + nva].tolist(), perm[ntr + nva:].tolist()
    tr = np.array([q in tr_q for q in Q])
    mu, sd = float(Y[tr].mean()), float(Y[tr].std() + 1e-6)
    model = OrderModel(ctx, seed=args.seed)
    t = time.perf_counter()
    losses, best, best_step, step = [], None, 0, 0
    while step < args.iters:
        n = min(args.eval_every, args.iters - step)
        losses += model.train(F[tr], (Y[tr] - mu) / sd, iters=n, lr=args.lr).tolist()
        step += n
        va = evaluate(model, F, Y, Q, C, mu, sd, va_q, args.margin)["learned_work"] if va_q else -step
        if best is None or va < best[0]:
            best = (va, model.get_params())
            best_step = step
    model.set_params(best[1])
    log(f"[train] {step} Adam steps on {int(tr.sum())} tuples in {time.perf_counter() - t:.2f}s, "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}; kept step {best_step} (least validation work)")
    ev = evaluate(model, F, Y, Q, C, mu, sd, te_q, args.margin)
    ev_train = evaluate(model, F, Y, Q, C, mu, sd, sorted(tr_q), args.margin)
    ev["no_margin"] = evaluate(model, F, Y, Q, C, mu, sd, te_q, 0.0)
    if va_q:
        ev["validation"] = evaluate(model, F, Y, Q, C, mu, sd, va_q, args.margin)
    ev["step"] = best_step
    log(f"[train] held-out: {json.dumps(ev)}")
    losses = np.array(losses, np.float32)
    # counts agree across all finished orders of each query
    consistent = all(len(set(CNT[(Q == q) & (CNT >= 0)].tolist())) <= 1 for q in range(len(qs)))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, params=model.get_params(), mu=mu, sd=sd, in_dim=GEM_FEAT_DIM, hidden=model.hidden,
             workload=args.workload, tuples=len(Y), queries=len(qs), losses=losses, margin=args.margin,
             heldout=json.dumps(ev), train=json.dumps(ev_train), counts_consistent=consistent)
    print(json.dumps({"tuples": int(len(Y)), "queries": len(qs), "heldout": ev, "train": ev_train,
                      "loss": [float(losses[0]), float(losses[-1])], "counts_consistent": bool(consistent),
                      "out": args.out}))
    ctx.close()


if __name__ == "__main__":
    main()

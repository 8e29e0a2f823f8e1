"""Train the order-cost MLP on tuples the matcher logs itself.

    python -m gem_amd.train_order_model --workload rmat22_learn --num-queries 768 \
        --out gem_amd/models/order_mlp_rmat22.npz
    python -m gem_amd.train_order_model --workload rmat20_train --num-queries 1280 --orders 16 \
        --out gem_amd/models/order_mlp_rmat20.npz

For every training query the library proposes candidate vertex orders with
their int32 features; the batched vertex-at-a-time matcher runs every (query,
order) pair at the benchmark's count cap (1e9) with a work limit, and logs the
rows of each level and the work done (gem_amd.learned.collect).  Labels are
log2(1 + work) relative to the query's cheapest order.  The MLP is fit on the
GPU (full-batch Adam, the library's own forward/backward kernels) on 70% of
the queries; the step count is the one whose rankings make the restart
schedule (gem_amd.learned.schedule_work, the rule the batched path applies)
spend the least work on another 15%; the last 15% are the held-out
evaluation: work of the schedule with the model's ranking against the
candidates' generation order (GEM's first) and against the best candidate of
each query.  The weights, label normalisation, schedule and evaluation are
saved to an .npz file; --cache keeps the collected tuples.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import GemContext
from .learned import OrderModel, collect, rank_orders, relative_targets, schedule_work
from .workloads import CAP, make_graph, make_queries


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def schedule_total(T, pred, queries, first, growth, width, limit, margin=0.0):
    """Work of the restart schedule over `queries`: ranked by `pred` (None =
    generation order), by the true costs ("best": each query's cheapest
    candidate first) or as given.  Queries none of whose candidates ended
    within the collection's work limit say nothing about the ranking and are
    left out."""
    tot = 0.0
    sel = np.isin(T["query"], [q for q in queries if not T["limited"][T["query"] == q].all()])
    Q, Cc, W = T["query"][sel], T["cand"][sel], T["work"][sel]
    lim = np.where(T["limited"][sel], np.inf, W)
    if isinstance(pred, str) and pred == "best":
        rk = {int(q): [int(c) for c in Cc[Q == q][np.argsort(lim[Q == q], kind="stable")]] for q in np.unique(Q)}
    elif pred is None:
        rk = {int(q): sorted(int(c) for c in Cc[Q == q]) for q in np.unique(Q)}
    else:
        rk = rank_orders(pred[sel], Q, Cc, margin)
    for q in np.unique(Q):
        m = Q == q
        costs = dict(zip(Cc[m].tolist(), lim[m].tolist()))
        tot += schedule_work(costs, rk[int(q)], first, growth, width)
    return tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="rmat22_learn")
    ap.add_argument("--num-queries", type=int, default=0, help="queries generated (0 = the workload's default)")
    ap.add_argument("--orders", type=int, default=8, help="candidate orders per query")
    ap.add_argument("--cap", type=float, default=CAP)
    ap.add_argument("--work-limit", type=float, default=1e10, help="work after which a (query, order) run stops")
    ap.add_argument("--group", type=int, default=64, help="(query, order) pairs per batched run")
    ap.add_argument("--cache", default=None, help=".npz of collected tuples: loaded if present, else written")
    ap.add_argument("--iters", type=int, default=1500, help="most Adam steps")
    ap.add_argument("--eval-every", type=int, default=50)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--first-attempt", type=float, default=1e8, help="restart schedule the model is evaluated with")
    ap.add_argument("--growth", type=int, default=4)
    ap.add_argument("--width", type=int, default=1)
    ap.add_argument("--margin", type=float, default=0.0,
                    help="log2 units by which a candidate must beat GEM's predicted cost to be ranked first")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "models", "order_mlp_rmat22.npz"))
    args = ap.parse_args()

    g = make_graph(args.workload)
    qs = make_queries(args.workload, g, args.num_queries or None)
    ctx = GemContext(0)
    ctx.load_data(g)
    if args.cache and os.path.exists(args.cache):
        z = np.load(args.cache, allow_pickle=False)
        T = {k: z[k] for k in z.files}
        log(f"[train] loaded {len(T['work'])} tuples from {args.cache}")
    else:
        t = time.perf_counter()
        T = collect(ctx, qs, args.orders, int(args.cap), int(args.work_limit), args.group, log=log)
        log(f"[train] {len(T['work'])} tuples from {len(qs)} queries in {time.perf_counter() - t:.0f}s")
        if args.cache:
            np.savez(args.cache, **T)
    lim = float(args.work_limit)
    # counts agree across all finished orders of each query
    consistent = all(len(set(T["count"][(T["query"] == q) & (T["count"] >= 0)].tolist())) <= 1
                     for q in np.unique(T["query"]))
    Y = relative_targets(np.minimum(T["work"], lim), T["query"])
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(len(qs))
    ntr, nva = int(0.7 * len(qs)), int(0.15 * len(qs))
    tr_q, va_q, te_q = set(perm[:ntr].tolist()), set(perm[ntr:ntr + nva].tolist()), set(perm[ntr + nva:].tolist())
    tr = np.isin(T["query"], list(tr_q))
    mu, sd = float(Y[tr].mean()), float(Y[tr].std() + 1e-6)
    sched = (args.first_attempt, args.growth, args.width, lim)
    model = OrderModel(ctx, seed=args.seed)
    t = time.perf_counter()
    losses, best, best_step, step = [], None, 0, 0
    while step < args.iters:
        n = min(args.eval_every, args.iters - step)
        losses += model.train(T["feats"][tr], (Y[tr] - mu) / sd, iters=n, lr=args.lr).tolist()
        step += n
        va = schedule_total(T, model.predict(T["feats"]) * sd + mu, va_q, *sched, margin=args.margin)
        if best is None or va < best[0]:
            best = (va, model.get_params())
            best_step = step
    model.set_params(best[1])
    log(f"[train] {step} Adam steps on {int(tr.sum())} tuples in {time.perf_counter() - t:.1f}s, loss "
        f"{losses[0]:.4f} -> {losses[-1]:.4f}; kept step {best_step} (least validation schedule work)")
    pred = model.predict(T["feats"]) * sd + mu

    def evaluate(sel):
        learned = schedule_total(T, pred, sel, *sched, margin=args.margin)
        gen = schedule_total(T, None, sel, *sched)
        bestw = schedule_total(T, "best", sel, *sched)
        unranked = sum(1 for q in sel if T["limited"][T["query"] == q].all())
        return {"queries": len(sel), "queries_no_order_ended": unranked, "learned_work": learned,
                "generation_order_work": gen, "best_first_work": bestw, "gain": gen / learned if learned else 0.0}

    ev = {"heldout": evaluate(te_q), "validation": evaluate(va_q), "train": evaluate(tr_q), "step": best_step,
          "schedule": {"first_attempt": args.first_attempt, "growth": args.growth, "width": args.width}}
    log(f"[train] {json.dumps(ev)}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, params=model.get_params(), mu=mu, sd=sd, in_dim=model.in_dim, hidden=model.hidden,
             workload=args.workload, tuples=len(Y), queries=len(qs), losses=np.array(losses, np.float32),
             margin=args.margin, cap=args.cap, work_limit=lim, heldout=json.dumps(ev["heldout"]),
             validation=json.dumps(ev["validation"]), train=json.dumps(ev["train"]), step=best_step,
             schedule=json.dumps(ev["schedule"]), counts_consistent=consistent)
    print(json.dumps({"tuples": int(len(Y)), "queries": len(qs), **ev, "loss": [losses[0], losses[-1]],
                      "counts_consistent": bool(consistent), "out": args.out}))
    ctx.close()


if __name__ == "__main__":
    main()

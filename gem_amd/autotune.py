"""Autotuner for the batched path.

    python -m gem_amd.autotune --workload rmat22 --out gem_amd/tuning/rmat22.json

Times the batched matcher on a sample of the workload's queries for every
combination of the tunables and keeps the fastest:
  * persistent grid size of the work-queue kernels (occupancy),
  * heavy-row split of the walk kernels (driving-list length from which a row
    is walked in pieces by many waves, and the piece length): load balance,
  * group size (queries resident on the GPU at once),
  * candidate orders scored per query by the order model.
The level budget (bytes of pending level chunks) is left automatic: a
quarter of the device memory free when a join starts; a sample of the
workload would not exercise it.  Counts do not depend on any of them
(checked on every trial).  The chosen configuration is stored as JSON and
read by bench.py.
"""
import argparse
import itertools
import json
import os
import sys
import time

from . import GemContext
from .learned import OrderModel, match_batch_learned
from .workloads import CAP, make_graph, make_queries

SPACE = {
    "persistent_blocks": [2048, 4096],
    "heavy_t": [256, 1024, 4096],
    "heavy_chunk": [1024, 4096],
    "group": [0],
    "max_orders": [8, 16],
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_tuning(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("best", {})
    return {}


def apply(ctx, cfg):
    """Set the context-level knobs of a tuning (the others are per-call
    arguments: level_budget, group, max_orders)."""
    if "persistent_blocks" in cfg:
        ctx.set_persistent_blocks(cfg["persistent_blocks"])
    if "heavy_t" in cfg:
        ctx.set_heavy_split(cfg["heavy_t"], cfg.get("heavy_chunk", 1024))
    if "max_active" in cfg:
        ctx.set_max_active(cfg["max_active"])
    if "graph" in cfg:
        ctx.set_graph(bool(cfg["graph"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="rmat22")
    ap.add_argument("--sample", type=int, default=96, help="queries timed per trial (every k-th of the workload)")
    ap.add_argument("--cap", type=float, default=CAP)
    ap.add_argument("--work-limit", type=float, default=5e9)
    ap.add_argument("--model", default=None, help="order model (default: models/order_mlp_<workload>.npz)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    out = args.out or os.path.join(os.path.dirname(__file__), "tuning", f"{args.workload}.json")

    g = make_graph(args.workload)
    allq = make_queries(args.workload, g)
    k = max(1, len(allq) // args.sample)
    qs = allq[::k][:args.sample]
    if args.model is None:
        args.model = os.path.join(os.path.dirname(__file__), "models", f"order_mlp_{args.workload}.npz")
    ctx = GemContext(0)
    ctx.load_data(g)
    model = OrderModel.load(ctx, args.model) if os.path.exists(args.model) else None
    cap, wl = int(args.cap), int(args.work_limit)
    match_batch_learned(ctx, qs, model, cap=cap, work_limit=wl)  # warm up
    trials, ref = [], None
    keys = list(SPACE)
    for vals in itertools.product(*(SPACE[k] for k in keys)):
        cfg = dict(zip(keys, vals))
        if model is None and cfg["max_orders"] != SPACE["max_orders"][0]:
            continue
        apply(ctx, cfg)
        t = time.perf_counter()
        res = match_batch_learned(ctx, qs, model, max_orders=cfg["max_orders"], cap=cap, work_limit=wl,
                                  group=cfg["group"])
        dt = time.perf_counter() - t
        sig = [(r["count"], r["saturated"], r["work_limited"]) for r in res]
        if ref is None:
            ref = sig
        same = all(a == b for a, b in zip(sig, ref) if not a[2] and not b[2])
        trials.append({**cfg, "seconds": dt, "counts_equal": same})
        log(f"[tune] {cfg} -> {dt:.3f}s counts_equal={same}")
        apply(ctx, {"persistent_blocks": 2048, "heavy_t": 1024, "heavy_chunk": 1024})
    ok = [t for t in trials if t["counts_equal"]]
    best = min(ok, key=lambda t: t["seconds"])
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    with open(out, "w") as f:
        json.dump({"workload": args.workload, "sample": len(qs), "best": {**{k: best[k] for k in keys}, "level_budget": 0},
                   "best_seconds": best["seconds"], "trials": trials}, f, indent=1)
    print(json.dumps({"best": {k: best[k] for k in keys}, "seconds": best["seconds"], "out": out}))
    ctx.close()


if __name__ == "__main__":
    main()

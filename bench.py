#!/usr/bin/env python
"""Benchmark: embeddings/s and queries/s of GEM subgraph matching on one MI355X.

Default workload: the 512-query R-MAT batch (2^22 vertices, 32 Zipf labels,
queries of 4-32 vertices).  A "step" is one pass of the batched matcher with
model-chosen join orders over the whole query set (filter -> position index ->
edge tables -> order candidates + MLP -> vertex-at-a-time join for every query,
many queries per kernel launch), with the data graph already resident in HBM.
`value` is that path's queries/s.  The plain path (the reference's per-query
pipeline with GEM's edge-at-a-time join) is timed in the same run over a
systematic sample of the queries, and the counts of the two are compared query
by query.  Counts are exact; on the R-MAT workloads each query's count is
capped at 1e9 (saturated queries report count = cap), and a query whose join
exceeds --work-limit is reported unfinished (plain path: rows read by its
steps; batched path: rows read plus list entries walked).

Prints ONE JSON line on stdout (progress goes to stderr).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "embeddings/s + queries/s, batched+learned-order vs plain GPU, count-exact"
HBM_PEAK_GBS = 8000.0

DESCR = {
    "er10k": "Erdos-Renyi 10,000 vertices / 50,000 edges, 8 uniform labels; 20 random-walk 5-vertex queries",
    "rmat20": "R-MAT 2^20 vertices, avg degree 16 (.57/.19/.19/.05), 16 Zipf labels; 100 random-walk 8-vertex "
              "queries (50 dense, 50 sparse); count capped at 1e9",
    "rmat22": "R-MAT 2^22 vertices, avg degree 16, 32 Zipf labels; 512 random-walk queries of 4-32 vertices; "
              "count capped at 1e9",
    "ooc24": "R-MAT 2^24 vertices, avg degree 64, 64 Zipf labels; 32 sparse random-walk 12-vertex queries; "
             "edge tables streamed from pinned host memory through a device cache; count capped at 1e9",
}


def parse_bytes(s):
    s = str(s).strip().upper()
    mult = {"K": 1 << 10, "M": 1 << 20, "G": 1 << 30}.get(s[-1:], 1)
    return int(float(s[:-1] if s[-1:] in "KMG" else s) * mult)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def profiled(key, kernel, workload):
    """A per-launch figure of `kernel` on `workload` recorded in the committed
    rocprofv3 summaries under profiles/ as a line `<key> <kernel> <workload> N`,
    or None when no summary covers it."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".md"):
            continue
        for line in open(os.path.join(pdir, fn)):
            f = line.split()
            if len(f) == 4 and f[0] == key and f[1] == kernel and f[2] == workload:
                return float(f[3])
    return None


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` (rocprofv3 FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy of known size), or None."""
    v = profiled("traffic_per_launch", kernel, workload)
    return int(v) if v is not None else None


def kernel_table(stats):
    return {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                "GB/s": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else 0.0}
            for k, v in stats.items() if v["launches"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2,
                    help="timed passes over the whole query set (the value is their mean rate)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat22", choices=sorted(DESCR))
    ap.add_argument("--queries", type=int, default=0, help="use only the first N queries")
    ap.add_argument("--cap", type=float, default=1e9)
    ap.add_argument("--work-limit", type=float, default=None,
                    help="work a query's join may do before it is reported unfinished (default: 5e9 on rmat22 "
                         "and ooc24, 2e10 on rmat20)")
    ap.add_argument("--group", type=int, default=None,
                    help="queries resident at once in the batched path (0 = all; default: autotuned, else 0)")
    ap.add_argument("--no-tuning", action="store_true", help="ignore gem_amd/tuning/<workload>.json")
    ap.add_argument("--level-budget", default=None,
                    help="bytes of pending level rows shared by the running queries (e.g. 64G; default: the "
                         "tuning's; 0 = a quarter of the device memory free when the batch starts)")
    ap.add_argument("--heavy-split", default=None,
                    help="THRESHOLD,CHUNK: rows with longer driving lists are walked in CHUNK-entry pieces "
                         "(default: autotuned, else 1024,1024)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path pass")
    ap.add_argument("--no-ablation", action="store_true", help="skip the pass with GEM-derived orders")
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="run the plain path on a systematic sample of this many queries (every k-th; "
                         "0 = all queries)")
    ap.add_argument("--model", default=None,
                    help="trained order model (.npz); default: the one trained for the workload's graph "
                         "(gem_amd/models/order_mlp_<graph>.npz); absent = GEM-derived orders")
    ap.add_argument("--orders", type=int, default=None,
                    help="candidate orders scored per query (default: autotuned, else 16)")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU-oracle work (0 = skip)")
    ap.add_argument("--cpu-work-limit", type=float, default=2e9,
                    help="intermediate rows after which one oracle query is abandoned")
    ap.add_argument("--stream-cache", default=None,
                    help="stream edge tables through a device cache of this size (e.g. 8G); default: 8G for ooc24, "
                         "tables in HBM otherwise")
    ap.add_argument("--stream-slot", default="1M", help="cache slot size for streamed tables")
    args = ap.parse_args()
    if args.stream_cache is None:
        args.stream_cache = "8G" if args.workload == "ooc24" else "0"
    stream_cache = parse_bytes(args.stream_cache)

    if args.gpus != 1:
        log("bench.py measures one GPU only: the multi-GPU form (root-partitioned join with one allreduce of "
            "the per-query counts) is not implemented, and this setup has a single MI355X per box.")
        sys.exit(2)

    from gem_amd import GemContext
    from gem_amd.workloads import make_graph, make_queries

    t0 = time.perf_counter()
    g = make_graph(args.workload)
    qs = make_queries(args.workload, g)
    if args.queries:
        qs = qs[:args.queries]
    cap = int(args.cap) if args.workload != "er10k" else 0
    if args.work_limit is None:
        args.work_limit = {"rmat20": 2e10, "er10k": 0}.get(args.workload, 5e9)
    wl = int(args.work_limit)
    log(f"[bench] {args.workload}: |V|={g.num_vertices} |E|={g.num_edges} queries={len(qs)} "
        f"dense={sum(q.dense for q in qs)} built in {time.perf_counter() - t0:.1f}s")

    ctx = GemContext(0)
    ctx.load_data(g)
    ctx.set_work_limit(wl)
    # the autotuner's choices for this workload, unless given on the command line
    from gem_amd.autotune import apply, load_tuning
    tuning = {} if args.no_tuning else load_tuning(os.path.join(ROOT, "gem_amd", "tuning", f"{args.workload}.json"))
    if args.heavy_split:
        ht, hc = (int(x) for x in args.heavy_split.split(","))
        tuning = {**tuning, "heavy_t": ht, "heavy_chunk": hc}
    apply(ctx, tuning)
    if args.group is None:
        args.group = int(tuning.get("group", 0))
    if args.orders is None:
        args.orders = int(tuning.get("max_orders", 16))
    level_budget = parse_bytes(args.level_budget) if args.level_budget else int(tuning.get("level_budget", 0))
    if tuning:
        log(f"[bench] autotuned: {tuning}")
    if stream_cache:
        ctx.set_table_stream(stream_cache, parse_bytes(args.stream_slot))
        log(f"[bench] edge tables streamed from host memory through a {args.stream_cache} device cache")

    # ---- batched path with the learned order (the measured value)
    from gem_amd.learned import OrderModel, match_batch_learned
    model = None
    if args.model is None:
        args.model = os.path.join(ROOT, "gem_amd", "models", f"order_mlp_{args.workload}.npz")
    if args.model and os.path.exists(args.model):
        model = OrderModel.load(ctx, args.model)
        log(f"[bench] order model {args.model}: trained on {model.meta.get('workload')} "
            f"({model.meta.get('tuples')} tuples)")
    else:
        log("[bench] no order model: batched path runs GEM-derived orders")

    def batched_pass():
        return match_batch_learned(ctx, qs, model, max_orders=args.orders, cap=cap, work_limit=wl,
                                   level_budget=level_budget,
                                   group=args.group)

    for w in range(args.warmup):
        t = time.perf_counter()
        batched_pass()
        log(f"[bench] batched warmup pass {w}: {time.perf_counter() - t:.2f}s")
    ctx.set_profiling(True)
    if stream_cache:
        ctx.stream_stats()  # reset
    t = time.perf_counter()
    batched = None
    emb = 0
    for s in range(args.steps):
        res = batched_pass()
        emb += sum(r["count"] for r in res)
        if batched is None:
            batched = res
        log(f"[bench] batched timed pass {s} done at {time.perf_counter() - t:.2f}s")
    elapsed = time.perf_counter() - t
    stats = ctx.kernel_stats()
    ctx.set_profiling(False)
    qps = len(qs) * args.steps / elapsed
    eps = emb / elapsed
    log("[bench] batched kernels:", json.dumps(kernel_table(stats)))

    # streamed tables: copy statistics of the timed passes, and the same
    # batched pass with the tables in HBM for the counts and the cost
    stream = None
    if stream_cache:
        ss = ctx.stream_stats()
        ctx.set_table_stream(0)
        t = time.perf_counter()
        res_h = batched_pass()
        th = time.perf_counter() - t
        ctx.set_table_stream(stream_cache, parse_bytes(args.stream_slot))
        stream = {"cache": args.stream_cache, "slot": args.stream_slot,
                  "table_bytes_per_pass": ss["table_bytes"] // args.steps,
                  "h2d_bytes_per_pass": ss["h2d_bytes"] // args.steps,
                  "h2d_GBps_during_copies": round(ss["h2d_bytes"] / (ss["copy_ms"] * 1e-3) / 1e9, 2)
                  if ss["copy_ms"] > 0 else None,
                  "copy_ms": round(ss["copy_ms"], 2), "streamed_step_ms": round(ss["step_ms"], 2),
                  "blocks_loaded": ss["blocks_loaded"], "blocks_hit": ss["blocks_hit"], "groups": ss["groups"],
                  "serial_groups": ss["serial_groups"],
                  "in_hbm_queries_per_s": round(len(qs) / th, 4),
                  "counts_equal_in_hbm": "{}/{} queries finished by both".format(
                      sum(1 for a, b in zip(batched, res_h) if not a["work_limited"] and not b["work_limited"]
                          and (a["count"], a["saturated"]) == (b["count"], b["saturated"])),
                      sum(1 for a, b in zip(batched, res_h) if not a["work_limited"] and not b["work_limited"]))}
        log(f"[bench] streamed tables: {stream}")

    # ablation: the same batched path with GEM-derived orders, one pass
    gem_batched = None
    if model is not None and not args.no_ablation:
        t = time.perf_counter()
        res_g = match_batch_learned(ctx, qs, None, cap=cap, work_limit=wl, level_budget=level_budget,
                                    group=args.group)
        tg = time.perf_counter() - t
        gem_batched = {"queries_per_s": round(len(qs) / tg, 4), "seconds": round(tg, 3),
                       "rows_processed": int(sum(r["rows_processed"] for r in res_g)),
                       "rows_processed_learned": int(sum(r["rows_processed"] for r in batched)),
                       "unfinished": sum(r["work_limited"] for r in res_g),
                       "counts_equal": sum(1 for a, b in zip(res_g, batched)
                                           if not a["work_limited"] and not b["work_limited"]
                                           and (a["count"], a["saturated"]) == (b["count"], b["saturated"]))}
        log(f"[bench] batched with GEM order: {gem_batched}")

    # roofline: the kernel with the most time in the timed region
    dom_name, dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9 if dom["ms"] > 0 else 0.0
    traffic = profiled_traffic(dom_name, args.workload)

    # ---- plain path, one pass over a systematic sample (every k-th query)
    plain = None
    if not args.no_plain:
        n = len(qs)
        k = max(1, -(-n // args.plain_sample)) if args.plain_sample else 1
        idx = list(range(0, n, k))
        sample = [qs[i] for i in idx]
        t = time.perf_counter()
        plain_res = []
        for j, q in enumerate(sample):
            plain_res.append(ctx.match_full(q, cap=cap))
            if (j + 1) % 10 == 0:
                log(f"[bench]   plain {j + 1}/{len(sample)} at {time.perf_counter() - t:.1f}s")
        pt = time.perf_counter() - t
        # the batched learned path over exactly the same queries, for a matched ratio
        t = time.perf_counter()
        matched = match_batch_learned(ctx, sample, model, max_orders=args.orders, cap=cap, work_limit=wl,
                                      level_budget=level_budget,
                                      group=args.group)
        mt = time.perf_counter() - t
        both = [(batched[i], b) for i, b in zip(idx, plain_res)
                if not batched[i]["work_limited"] and not b["work_limited"]]
        agree = sum(1 for a, b in both if (a["count"], a["saturated"]) == (b["count"], b["saturated"]))
        plain = {"queries_per_s": round(len(sample) / pt, 4), "embeddings_per_s": round(
            sum(r["count"] for r in plain_res) / pt, 1), "seconds": round(pt, 3),
                 "sample": f"every {k}-th query: {len(sample)} of {n}",
                 "unfinished": sum(r["work_limited"] for r in plain_res),
                 "batched_on_sample_queries_per_s": round(len(sample) / mt, 4),
                 "counts_equal_to_batched": f"{agree}/{len(both)} queries finished by both"}
        log(f"[bench] plain: {plain}")

    # ---- CPU baseline: the oracle on a bounded prefix of the same queries
    cpu = None
    if args.cpu_budget > 0:
        import numpy as np

        import oracle
        from gem_amd.graph import gem_order
        # the same systematic sample as the plain path, in order, until the
        # budget is spent; the edge order is GEM's (from the oracle's own table
        # sizes, outside the timed call: the lattice form is exponential in the
        # edges of the larger queries)
        n = len(qs)
        k = max(1, -(-n // args.plain_sample)) if args.plain_sample else 1
        spent, done, abandoned, mism = 0.0, 0, 0, 0
        for i in range(0, n, k):
            q = qs[i]
            if spent >= args.cpu_budget:
                break
            sizes = oracle.table_sizes(g, q)
            order = gem_order(q, np.maximum(sizes[0::2], sizes[1::2]))
            tq = time.perf_counter()
            r = oracle.count(g, q, order=order, cap=cap, work_limit=int(args.cpu_work_limit))
            spent += time.perf_counter() - tq
            if r is None:
                abandoned += 1
                continue
            done += 1
            gr = batched[i]
            if not gr["work_limited"] and (r["count"], r["saturated"]) != (gr["count"], gr["saturated"]):
                mism += 1
        threads = oracle.max_threads()
        cpu = {"value": round(done / spent, 4) if spent > 0 else 0.0, "unit": "queries/s", "cores": threads,
               "kind": "port",
               "sample": f"every {k}-th query until {args.cpu_budget:.0f}s: {done + abandoned} of {n} queries "
                         f"({done} finished, {abandoned} abandoned after {args.cpu_work_limit:.0e} intermediate "
                         f"rows), {spent:.1f}s on {threads} threads of {os.cpu_count()} ({cpu_model()}); counts "
                         f"equal to GPU on {done - mism}/{done}"}
        log(f"[bench] cpu oracle: {cpu}")

    out = {
        "metric": METRIC,
        "value": round(qps, 4),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded generators)",
        "config": {"workload": args.workload, "description": DESCR[args.workload], "queries": len(qs),
                   "path": "batched vertex-at-a-time join, " + (
                       "learned order (MLP over candidate orders)" if model is not None else "GEM-derived order"),
                   "cap": cap,
                   "work_limit": wl,
                   "tuning": tuning or None},
        "embeddings_per_s": round(eps, 1),
        "saturated_queries": sum(r["saturated"] for r in batched),
        "unfinished_queries": sum(r["work_limited"] for r in batched),
        "plain": plain,
        "batched_gem_order": gem_batched,
        "speedup_vs_plain": round(qps / plain["queries_per_s"], 3) if plain else None,
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": round(dom["bytes"] / max(dom["launches"], 1)),
                     "launches": dom["launches"], "avg_launch_ms": round(dom["ms"] / max(dom["launches"], 1), 3),
                     "profiled_avg_launch_ms": (round(prof_us / 1000, 3) if (prof_us := profiled(
                         "avg_launch_us", dom_name, args.workload)) is not None else None)},
        "cpu_baseline": cpu,
        "streamed_tables": stream,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

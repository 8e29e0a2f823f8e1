#!/usr/bin/env python
"""Benchmark: embeddings/s and queries/s of GEM subgraph matching on one MI355X.

Default workload: the 512-query R-MAT batch (2^22 vertices, 32 Zipf labels,
queries of 4-32 vertices).  A "step" is one pass of the batched matcher with
model-ranked join orders over the whole query set (filter -> position index
-> edge tables -> order candidates + MLP -> vertex-at-a-time join for every
query, many queries per kernel launch), with the data graph already resident
in HBM.  `value` is that path's queries/s.  Every query ends with its exact
count or, above the cap (1e9 on the R-MAT workloads), saturated: a query
whose attempt does more work than its limit starts again with the next
ranked order and twice the limit (the count does not depend on the order).
The timed passes run without per-launch profiling (hipGraph replay of the
extend step); one more pass with profiling on gives the kernel table and the
roofline.  The plain path (the reference's per-query pipeline with GEM's
edge-at-a-time join) runs in the same process over a systematic sample of
the queries with a work limit of its own (rows read); queries it does not
finish make its rate, and the speedup over it, a bound.  Counts are compared
query by query with the plain path and with the CPU oracle.

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
    "rmat20_learn": "R-MAT 2^20 vertices, avg degree 16, 16 Zipf labels; 256 random-walk queries of 8-24 vertices "
                    "(learned-order workload); count capped at 1e9",
    "ooc24": "R-MAT 2^24 vertices, avg degree 64, 64 Zipf labels; 32 sparse random-walk 12-vertex queries; "
             "data graph's neighbours in pinned host memory, edge tables streamed from pinned host memory "
             "through a device cache; count capped at 1e9",
    "ooc26": "R-MAT 2^26 vertices, avg degree 64, 64 Zipf labels; 32 sparse random-walk 12-vertex queries; "
             "data graph's neighbours in pinned host memory, edge tables streamed from pinned host memory "
             "through a device cache; count capped at 1e9",
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
    """A figure of `kernel` on `workload` (per launch or per pass) recorded in
    the committed rocprofv3 summaries under profiles/ as a line
    `<key> <kernel> <workload> N`, or None when no summary covers it."""
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


def digest(res):
    """sha256 (16 hex digits) of the per-query (count, saturated) list."""
    import hashlib
    return hashlib.sha256(json.dumps([[int(r["count"]), bool(r["saturated"])] for r in res]).encode()).hexdigest()[:16]


def mem_line(ctx):
    m = ctx.memory_stats(reset_peak=True)
    gib = 1 << 30
    return (f"pool high-water {m['peak_in_use'] / gib:.1f} GiB (held {m['held'] / gib:.1f} GiB, in use after "
            f"{m['in_use'] / gib:.2f} GiB); level budget {m['level_budget'] / gib:.1f} GiB, peak charged "
            f"{m['peak_charged'] / gib:.1f} GiB, {m['iterations']} iterations, {m['forced']} forced")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2,
                    help="timed passes over the whole query set (the value is their mean rate)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat22", choices=sorted(DESCR))
    ap.add_argument("--queries", type=int, default=0, help="use only the first N queries")
    ap.add_argument("--cap", type=float, default=1e9)
    ap.add_argument("--first-attempt", type=float, default=1e8,
                    help="work (rows read + list entries walked) of a query's first attempt in the batched path; "
                         "attempt a runs its (a mod n)-th ranked order and may do first*growth^(a // width) before "
                         "the query starts again (0 = no restarts)")
    ap.add_argument("--restart-growth", type=int, default=4)
    ap.add_argument("--restart-width", type=int, default=1)
    ap.add_argument("--work-limit", type=float, default=1e12,
                    help="total work over a batched query's attempts after which it is reported unfinished (0 = "
                         "none; the default is a safety net about a hundred times what the costliest query of the "
                         "512-query batch needs)")
    ap.add_argument("--plain-work-limit", type=float, default=5e9,
                    help="rows the plain path's steps may read per query before it is reported unfinished")
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
    ap.add_argument("--no-ablation", action="store_true", help="skip the pass without the order model")
    ap.add_argument("--no-profile-pass", action="store_true", help="skip the profiled pass (no roofline)")
    ap.add_argument("--plain-sample", type=int, default=32,
                    help="run the plain path on a systematic sample of this many queries (every k-th; "
                         "0 = all queries)")
    ap.add_argument("--model", default=None,
                    help="trained order model (.npz); default: the one trained for the workload's graph "
                         "(gem_amd/models/order_mlp_<graph>.npz); absent = candidates in generation order")
    ap.add_argument("--orders", type=int, default=None,
                    help="candidate orders ranked per query (default: autotuned, else 8)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-oracle work (0 = skip)")
    ap.add_argument("--cpu-work-limit", type=float, default=2e9,
                    help="intermediate rows after which one oracle query is abandoned")
    ap.add_argument("--stream-cache", default=None,
                    help="stream edge tables through a device cache of this size (e.g. 8G); default: 8G for ooc24 / ooc26, "
                         "tables in HBM otherwise")
    ap.add_argument("--stream-slot", default="1M", help="cache slot size for streamed tables")
    ap.add_argument("--no-racing", action="store_true", help="no twins racing the batch's stragglers")
    ap.add_argument("--no-prep-overlap", action="store_true",
                    help="prepare every query before the join starts (no second stream)")
    ap.add_argument("--graph-host", choices=["auto", "on", "off"], default="auto",
                    help="keep the data graph's neighbour array in pinned host memory (auto: on for ooc24)")
    args = ap.parse_args()
    graph_host = args.graph_host == "on" or (args.graph_host == "auto" and args.workload.startswith("ooc"))
    if args.stream_cache is None:
        args.stream_cache = "8G" if args.workload.startswith("ooc") else "0"
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
    wl = int(args.work_limit)
    first = int(args.first_attempt)
    log(f"[bench] {args.workload}: |V|={g.num_vertices} |E|={g.num_edges} queries={len(qs)} "
        f"dense={sum(q.dense for q in qs)} built in {time.perf_counter() - t0:.1f}s")

    ctx = GemContext(0)
    if graph_host:
        ctx.set_graph_residency(True)
    t = time.perf_counter()
    ctx.load_data(g)
    log(f"[bench] data graph loaded in {time.perf_counter() - t:.1f}s, neighbour array in "
        f"{'pinned host memory (read over PCIe)' if graph_host else 'HBM'}")
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
        args.orders = int(tuning.get("max_orders", 8))
    level_budget = parse_bytes(args.level_budget) if args.level_budget else int(tuning.get("level_budget", 0))
    if tuning:
        log(f"[bench] autotuned: {tuning}")
    if stream_cache:
        ctx.set_table_stream(stream_cache, parse_bytes(args.stream_slot))
        log(f"[bench] edge tables streamed from host memory through a {args.stream_cache} device cache")
    ctx.set_racing(not args.no_racing)
    ctx.set_prep_overlap(not args.no_prep_overlap)

    # ---- batched path with the learned order (the measured value)
    from gem_amd.learned import OrderModel, match_batch_learned
    model = None
    if args.model is None:
        graph_name = {"rmat20_learn": "rmat20"}.get(args.workload, args.workload)
        args.model = os.path.join(ROOT, "gem_amd", "models", f"order_mlp_{graph_name}.npz")
    if args.model and os.path.exists(args.model):
        model = OrderModel.load(ctx, args.model)
        log(f"[bench] order model {args.model}: trained on {model.meta.get('workload')} "
            f"({model.meta.get('tuples')} tuples)")
    else:
        log("[bench] no order model: candidate orders run in generation order (GEM's first)")

    def batched_pass(m=model, queries=qs):
        return match_batch_learned(ctx, queries, m, max_orders=args.orders, cap=cap, work_limit=wl,
                                   level_budget=level_budget, group=args.group, first_attempt=first,
                                   restart_growth=args.restart_growth, restart_width=args.restart_width)

    ctx.memory_stats(reset_peak=True)
    for w in range(args.warmup):
        t = time.perf_counter()
        batched_pass()
        log(f"[bench] batched warmup pass {w}: {time.perf_counter() - t:.2f}s; {mem_line(ctx)}")
    if stream_cache:
        ctx.stream_stats()  # reset
    # timed passes: no per-launch profiling (the extend step replays its hipGraph)
    ctx.set_profiling(False)
    pass_s, digests, prep = [], [], []
    batched = None
    emb = 0
    for s in range(args.steps):
        t = time.perf_counter()
        res = batched_pass()
        pass_s.append(time.perf_counter() - t)
        pt = ctx.phase_times()
        prep.append((pt["filter_ms"], pt["tables_ms"]))
        emb += sum(r["count"] for r in res)
        digests.append(digest(res))
        if batched is None:
            batched = res
        log(f"[bench] batched timed pass {s}: {pass_s[-1]:.2f}s, digest {digests[-1]}, unfinished "
            f"{sum(r['work_limited'] for r in res)}, restarts {sum(r['restarts'] for r in res)}; {mem_line(ctx)}")
    elapsed = sum(pass_s)
    qps = len(qs) * args.steps / elapsed
    eps = emb / elapsed

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
                  "counts_equal_in_hbm": digest(res_h) == digests[0]}
        log(f"[bench] streamed tables: {stream}")

    # one more pass with per-launch profiling: kernel table and roofline
    stats, prof_s = {}, None
    if not args.no_profile_pass:
        ctx.set_profiling(True)
        t = time.perf_counter()
        res_p = batched_pass()
        prof_s = time.perf_counter() - t
        stats = ctx.kernel_stats()
        ctx.set_profiling(False)
        digests.append(digest(res_p))
        log(f"[bench] profiled pass: {prof_s:.2f}s; kernels: {json.dumps(kernel_table(stats))}")

    # ablation: the same batched path without the order model (candidates in
    # generation order, GEM's first, same restart rule), one pass
    no_model = None
    if model is not None and not args.no_ablation:
        t = time.perf_counter()
        res_g = batched_pass(m=None)
        tg = time.perf_counter() - t
        no_model = {"queries_per_s": round(len(qs) / tg, 4), "seconds": round(tg, 3),
                    "work": int(sum(r["work"] for r in res_g)),
                    "work_learned": int(sum(r["work"] for r in batched)),
                    "rows_processed": int(sum(r["rows_processed"] for r in res_g)),
                    "rows_processed_learned": int(sum(r["rows_processed"] for r in batched)),
                    "restarts": sum(r["restarts"] for r in res_g),
                    "unfinished": sum(r["work_limited"] for r in res_g),
                    "counts_equal": digest(res_g) == digests[0]}
        log(f"[bench] batched without the order model: {no_model}")

    # roofline: the kernel with the most time in the profiled pass
    roof = None
    if stats:
        dom_name, dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
        achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9 if dom["ms"] > 0 else 0.0
        prof_us = profiled("avg_launch_us", dom_name, args.workload)
        prof_total = profiled("total_us_per_pass", dom_name, args.workload)
        prof_launches = profiled("launches_per_pass", dom_name, args.workload)
        roof = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(dom_name, args.workload),
                "algorithmic_bytes_per_launch": round(dom["bytes"] / max(dom["launches"], 1)),
                "launches": dom["launches"], "avg_launch_ms": round(dom["ms"] / max(dom["launches"], 1), 3),
                "profiled_avg_launch_ms": round(prof_us / 1000, 3) if prof_us is not None else None,
                # per pass: launches follow the pass's scheduler iterations,
                # which vary from pass to pass, and so does the time per launch
                "ms_per_pass": round(dom["ms"], 1),
                "profiled_ms_per_pass": round(prof_total / 1000, 1) if prof_total is not None else None,
                "profiled_launches": int(prof_launches) if prof_launches is not None else None}

    # ---- plain path, one pass over a systematic sample (every k-th query)
    plain = None
    n = len(qs)
    k = max(1, -(-n // args.plain_sample)) if args.plain_sample else 1
    idx = list(range(0, n, k))
    if not args.no_plain:
        sample = [qs[i] for i in idx]
        ctx.set_work_limit(int(args.plain_work_limit))
        t = time.perf_counter()
        plain_res = []
        for j, q in enumerate(sample):
            plain_res.append(ctx.match_full(q, cap=cap))
            if (j + 1) % 8 == 0:
                log(f"[bench]   plain {j + 1}/{len(sample)} at {time.perf_counter() - t:.1f}s")
        pt = time.perf_counter() - t
        ctx.set_work_limit(0)
        # the batched learned path over exactly the same queries, for a matched ratio
        t = time.perf_counter()
        batched_pass(queries=sample)
        mt = time.perf_counter() - t
        both = [(batched[i], b) for i, b in zip(idx, plain_res) if not b["work_limited"]]
        agree = sum(1 for a, b in both if (a["count"], a["saturated"]) == (b["count"], b["saturated"]))
        unf = sum(r["work_limited"] for r in plain_res)
        plain = {"queries_per_s": round(len(sample) / pt, 4), "embeddings_per_s": round(
            sum(r["count"] for r in plain_res) / pt, 1), "seconds": round(pt, 3),
                 "sample": f"every {k}-th query: {len(sample)} of {n}",
                 "work_limit_rows": int(args.plain_work_limit),
                 "unfinished": unf,
                 "rate_is": "an upper bound (unfinished queries stopped early)" if unf else "exact",
                 "batched_on_sample_queries_per_s": round(len(sample) / mt, 4),
                 "counts_equal_to_batched": f"{agree}/{len(both)} queries the plain path finished"}
        log(f"[bench] plain: {plain}")

    # ---- CPU baseline: the oracle on a bounded prefix of the same sample
    cpu = None
    if args.cpu_budget > 0:
        import numpy as np

        import oracle
        from gem_amd.graph import gem_order
        # the same systematic sample as the plain path, in order, until the
        # budget is spent; the edge order is GEM's (from the oracle's own table
        # sizes, outside the timed call: the lattice form is exponential in the
        # edges of the larger queries)
        spent, done, abandoned, mism = 0.0, 0, 0, 0
        for i in idx:
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
            if (r["count"], r["saturated"]) != (gr["count"], gr["saturated"]):
                mism += 1
        threads = oracle.max_threads()
        cpu = {"value": round(done / spent, 4) if spent > 0 else 0.0, "unit": "queries/s", "cores": threads,
               "kind": "port",
               "sample": f"every {k}-th query until {args.cpu_budget:.0f}s: {done + abandoned} of {n} queries "
                         f"({done} finished, {abandoned} abandoned after {args.cpu_work_limit:.0e} intermediate "
                         f"rows), {spent:.1f}s on {threads} threads of {os.cpu_count()} ({cpu_model()}); counts "
                         f"equal to GPU on {done - mism}/{done}"}
        log(f"[bench] cpu oracle: {cpu}")

    plain_q = plain["queries_per_s"] if plain else None
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
                       "learned order (MLP-ranked candidate orders)" if model is not None
                       else "candidate orders in generation order"),
                   "cap": cap, "restarts": {"first_attempt_work": first, "growth": args.restart_growth,
                                            "width": args.restart_width,
                                            "racing_twins": not args.no_racing and not stream_cache},
                   "prep_overlap": not args.no_prep_overlap and not stream_cache,
                   "work_limit": wl,
                   "tuning": tuning or None},
        "embeddings_per_s": round(eps, 1),
        "pass_seconds": [round(x, 3) for x in pass_s],
        "filter_tables_seconds_per_pass": [round(sum(p[0] for p in prep) / len(prep) / 1e3, 3),
                                           round(sum(p[1] for p in prep) / len(prep) / 1e3, 3)],
        "graph_residency": "host (pinned, read over PCIe)" if graph_host else "HBM",
        "saturated_queries": sum(r["saturated"] for r in batched),
        "exact_queries": sum(not r["saturated"] and not r["work_limited"] for r in batched),
        "unfinished_queries": sum(r["work_limited"] for r in batched),
        "restarts": sum(r["restarts"] for r in batched),
        "counts_digest": digests[0],
        "digests_agree": len(set(digests)) == 1,
        "no_model": no_model,
        "learned_gain": round(no_model["seconds"] / (elapsed / args.steps), 3) if no_model else None,
        "plain": plain,
        "speedup_vs_plain": round(qps / plain_q, 3) if plain_q else None,
        "speedup_is": ("a lower bound: the plain rate counts queries it stopped at its work limit"
                       if plain and plain["unfinished"] else "exact") if plain else None,
        "roofline": roof,
        "profiled_pass_seconds": round(prof_s, 3) if prof_s is not None else None,
        "cpu_baseline": cpu,
        "streamed_tables": stream,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

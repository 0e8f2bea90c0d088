#!/usr/bin/env python3
"""Benchmark of the rule support / body-pair counting path on one MI355X.

Workload (default): the DBpedia-shaped configuration -- 206,000 triples over
50,000 entities and 300 predicates (predicate frequency Zipf s=1.0,
per-predicate out-degree lognormal(0,1)), 10,000 closed rules of 1-3 body
atoms -- generated from a fixed seed (the reference's datasets are not
available, so the data is synthetic).

A step is one evaluation of all rules: exact support and distinct body-pair
count for every rule (the body pairs of every rule are counted, also when
support is 0, i.e. the work CPU_01_nonparallel does), from a batch compiled
once and kept resident in HBM, replayed as a HIP graph. The JSON line also
reports:
  * per_rule: the same rules through the per-rule launch sequence of the
    reference library (one support launch and one body-size launch per rule,
    each followed by a sync and a device-to-host copy), with p50/p99 latency;
  * roofline: algorithmic bytes of the dominant kernel (host traffic model of
    the same plans) over its event-timed duration, against 8 TB/s;
  * cpu_baseline: the reference's CPU_01_nonparallel program (built from the
    reference sources into oracle/_ref/, single-threaded) on a bounded sample
    of the same rules, or the project's CPU oracle if that build is absent.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from masterthesis_gpu_acceleration_amd import lib as L  # noqa: E402
from masterthesis_gpu_acceleration_amd import workloads as W  # noqa: E402

METRIC = "rules/s & body-pairs/s (batched vs per-rule launch); p50/p99 rule latency"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(x, q):
    x = np.sort(np.asarray(x, np.float64))
    if x.size == 0:
        return 0.0
    k = max(0, min(x.size - 1, int(np.ceil(q * x.size)) - 1))
    return float(x[k])


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


WORKLOADS = {
    # name: (generator, CPU-baseline sample size in rules)
    "dbpedia_shaped": (lambda: W.dbpedia_shaped(), 2000),
    "small_paths": (lambda: W.small_paths(), 10),
    "hub_fanout": (lambda: W.hub_fanout(), 3),
    "dense_conf": (lambda: W.dense_conf(), 8),
}


def pmc_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, scaled by the copy yardstick),
    or None when that workload has not been profiled."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        d = json.load(open(p))[workload]
        return int(d["kernels"][kernel]["hbm_bytes"]), d["kernels"][kernel]["l2_hit"], d["source"]
    except (OSError, KeyError, ValueError):
        return None, None, None


def snapshot_leg(eng, words, flags, res, load_triples_s):
    """Graph load times: from term-id triples (host sort + index build +
    upload) against a binary snapshot of the same graph (map + checks +
    upload); the snapshot's context must give the same counts."""
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.snap")
        t = time.time()
        eng.save_snapshot(p)
        save_s = time.time() - t
        size = os.path.getsize(p)
        t = time.time()
        e2 = L.Engine.from_snapshot(p)
        snap_s = time.time() - t
        try:
            r = e2.evaluate(words, flags)
            same = bool(np.array_equal(r["support"], res["support"]) and np.array_equal(r["bodyPairs"], res["bodyPairs"]))
        finally:
            e2.close()
    log("load: from triples %.2fs, snapshot save %.2fs (%.1f MB), from snapshot %.2fs, same counts %s" % (
        load_triples_s, save_s, size / 1e6, snap_s, same))
    return {"from_triples": round(load_triples_s, 3), "snapshot_save": round(save_s, 3),
            "from_snapshot": round(snap_s, 3), "snapshot_mb": round(size / 1e6, 1), "same_counts": same}


def describe(wl):
    g = wl.graph
    nb = [len(r[1]) for r in wl.rules]
    return "%s: %d triples, %d entities, %d predicates, %d rules of %d-%d atoms" % (
        wl.name, g.n_triples, g.n_entities, g.n_predicates, len(wl.rules), min(nb), max(nb))


def cpu_baseline(wl, rules, sample):
    """Reference CPU_01_nonparallel (single thread) on the first `sample`
    rules; rule compute = sum of its per-rule support + confidence times, as
    the reference notebooks count it (RunBenchmark.ipynb; thesis p.43)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cpu01")
    sub = rules[:sample]
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as d:
            ttl = os.path.join(d, "g.ttl")
            rf = os.path.join(d, "r.txt")
            W.write_ttl(wl.graph, ttl)
            W.write_rules(sub, rf)
            t = time.time()
            try:
                out = subprocess.run([exe, ttl, rf, "--json"], capture_output=True, text=True, timeout=180,
                                     check=True)
            except subprocess.TimeoutExpired:
                return {"value": None, "unit": "rules/s", "cores": 1, "kind": "reference",
                        "sample": "first %d rules did not finish within 180 s on one core" % len(sub)}
            wall = time.time() - t
        rows = json.loads(out.stdout)
        ms = sum(r["supportMs"] + r["confidenceMs"] for r in rows)
        return {"value": round(len(rows) / (ms / 1000.0), 2), "unit": "rules/s", "cores": 1, "kind": "reference",
                "sample": "first %d of the %d rules, full graph; CPU_01_nonparallel rule compute %.1f ms "
                          "(wall incl. Turtle load %.1f s); host %s, nproc %d" % (
                              len(rows), len(rules), ms, wall, cpu_model(), os.cpu_count() or 0)}
    from oracle import Oracle  # CPU oracle (the port) when the reference build is absent
    g = wl.graph
    o = Oracle(g.s, g.p, g.o)
    words = W.encode_rules(sub)
    t = time.time()
    o.eval(words)
    dt = time.time() - t
    return {"value": round(len(sub) / dt, 2), "unit": "rules/s", "cores": 1, "kind": "port",
            "sample": "first %d of the %d rules, full graph; project CPU oracle; host %s, nproc %d" % (
                len(sub), len(rules), cpu_model(), os.cpu_count() or 0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dbpedia_shaped", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=int, default=0, help="rules in the CPU-baseline sample (0: per workload)")
    ap.add_argument("--gated", action="store_true", help="body pairs only for rules with support > 0")
    ap.add_argument("--serve", action="store_true",
                    help="serving mode: Poisson request stream over the high-fan-out graph")
    ap.add_argument("--requests", type=int, default=6000, help="requests in the serving stream")
    ap.add_argument("--rate", type=float, default=150.0, help="requests per second (serving mode)")
    ap.add_argument("--rates", default="75,150,300,600,1200",
                    help="serving mode: arrival rates of the rate sweep (comma list; empty: no sweep)")
    ap.add_argument("--sweep-requests", type=int, default=1500, help="requests per rate of the serving sweep")
    ap.add_argument("--max-work", type=float, default=1e8,
                    help="serving: drop requests whose estimated search work exceeds this (0: keep all)")
    ap.add_argument("--compare-dedup", action="store_true",
                    help="also time the batch under each body-pair dedup strategy (default for dense_conf)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-snapshot", action="store_true", help="skip the graph-load leg (triples vs snapshot)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="graph replays run the support and body-pair phases one after the other (per-kernel profiles)")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="extra tuning key for the engine (gpu_set_tuning), applied before autotuning")
    ap.add_argument("--no-fuse", action="store_true",
                    help="every rule's support gets its own pass (no support counting in the body-pair pass)")
    ap.add_argument("--no-per-rule", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune", default="",
                    help="BUDGET,THREAD_WAVES_PER_CU,WAVE_WAVES_PER_CU[,DEDUP[,GIANT_AFTER[,CHUNKS_PER_WAVE]]]: fixed "
                         "tuning, no autotune")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the streaming copy yardstick over this many bytes (counter scaling) and exit")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path (rules are independent: "
              "each GPU would hold the whole graph and take a share of the rules). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        g = W.tiny_random(1)
        eng = L.Engine.from_graph(g)
        ms = eng.lib.gpu_copy_probe(eng.ctx, int(args.copy_probe) // 16 * 16)
        print(json.dumps({"copy_probe_bytes": int(args.copy_probe) // 16 * 16, "ms": ms,
                          "GBps": (2 * args.copy_probe / (ms * 1e-3) / 1e9) if ms > 0 else None}), flush=True)
        eng.close()
        return

    if args.serve:
        return serve_mode(args)

    t0 = time.time()
    gen, sample = WORKLOADS[args.workload]
    wl = gen()
    g = wl.graph
    rules = wl.rules
    log("workload %s (%.1fs)" % (describe(wl), time.time() - t0))
    words = L.rule_words_for_graph(rules, g)
    t = time.time()
    eng = L.Engine.from_graph(g)
    load_triples_s = time.time() - t
    log("graph resident in HBM (%.2fs)" % load_triples_s)

    flags = 0 if args.gated else L.UNGATED
    if args.no_overlap:
        eng.set_tuning("overlap", 0)
    if args.no_fuse:
        eng.set_tuning("fuse_support", 0)
    for kv in args.set:
        key, val = kv.split("=", 1)
        eng.set_tuning(key, int(val))
    if args.tune:
        knobs = [int(x) for x in args.tune.split(",")]
        for key, val in zip(("budget", "thread_waves_per_cu", "wave_waves_per_cu", "dedup", "giant_after",
                             "chunks_per_wave"), knobs):
            eng.set_tuning(key, val)
        log("fixed tuning: %s" % args.tune)
    elif not args.no_autotune:
        t = time.time()
        tl = eng.autotune(words, flags)
        best = tl[np.argmin(tl[:, 3])]
        log("autotune (%.1fs, %d trials): budget %d, thread waves/CU %d, wave waves/CU %d -> %.3f ms; "
            "chunks/wave %d, giant_after %d" % (time.time() - t, len(tl), best[0], best[1], best[2], best[3],
                                                eng.get_tuning("chunks_per_wave"), eng.get_tuning("giant_after")))
        for r in tl:
            log("   trial budget %6d tw %2d ww %2d  %.3f ms  escalated %d  pairs dedup %d  giant_after %d" % tuple(r))
    t = time.time()
    batch = eng.batch(words, flags)
    build_s = time.time() - t
    log("batch compiled and uploaded (%.3fs)" % build_s)
    for _ in range(args.warmup):
        res = batch.run()
    log("warmup done")
    dev_ms = []
    t = time.perf_counter()
    for _ in range(args.steps):
        res = batch.run()
        dev_ms.append(eng.last_stats()["total_ms"])
    wall = time.perf_counter() - t
    ms_step = wall * 1000.0 / args.steps
    res = res.copy()
    log("timed %d steps: %.3f ms/step (device %.3f ms)" % (args.steps, ms_step, float(np.mean(dev_ms))))
    n = len(rules)
    pairs = int(res["bodyPairs"][res["bodyPairs"] > 0].sum())
    rules_per_s = n / (ms_step / 1000.0)

    # end to end: compile + upload + run of a fresh batch
    t = time.perf_counter()
    eng.evaluate(words, flags)
    e2e_s = time.perf_counter() - t

    # roofline of the dominant kernel: eager pass with per-kernel events
    eng.set_tuning("use_graph", 0)
    kms = np.zeros(4)
    reps = 3
    for _ in range(reps):
        eng.evaluate(words, flags)
        kms += eng.kernel_ms()
    kms /= reps
    eng.set_tuning("use_graph", 1)
    traffic = eng.traffic(words)
    names = ["k_thread_batch<0>", "k_wave_batch<0>", "k_thread_batch<1>", "k_wave_batch<1>"]   # <plan kind>
    dom = int(np.argmax(kms))
    phase = dom // 2
    bytes_phase = int(traffic[:, phase].sum())
    achieved = bytes_phase / (kms[dom] * 1e-3) / 1e9 if kms[dom] > 0 else 0.0
    heavy = eng.last_stats()["heavy_seeds"]
    hbm, l2hit, src = pmc_traffic(args.workload + ("_gated" if args.gated else ""), names[dom])
    log("kernels ms: %s; escalated seeds %d" % (", ".join("%s=%.3f" % (a, b) for a, b in zip(names, kms)), heavy))

    out = {
        "metric": METRIC,
        "value": round(rules_per_s, 1),
        "unit": "rules/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": describe(wl) + (", exact support + distinct body pairs for %s, resident batch, "
                                                "HIP graph replay" % ("rules with support > 0" if args.gated
                                                                      else "every rule")),
                   "rules": n, "triples": g.n_triples},
        "body_pairs_per_s": round(pairs / (ms_step / 1000.0), 1),
        "body_pairs_total": pairs,
        "rules_with_support": int((res["support"] > 0).sum()),
        "escalated_seeds": int(heavy),
        "tuning": {k: eng.get_tuning(k) for k in ("budget", "thread_waves_per_cu", "wave_waves_per_cu", "pairs_dedup",
                                                  "giant_after", "chunks_per_wave", "wave_slots", "fuse_active")},
        "support_fused_rules": eng.get_tuning("last_fused"),
        "e2e_rules_per_s": round(n / e2e_s, 1),
        "batch_build_s": round(build_s, 4),
        "latency_ms": {"batched_p50": round(pct([x for x in dev_ms], 0.5), 4),
                       "batched_p99": round(pct([x for x in dev_ms], 0.99), 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": hbm,
                     "traffic_source": src, "l2_hit": l2hit, "kernel": names[dom], "kernel_ms": round(float(kms[dom]), 4),
                     "algorithmic_bytes": bytes_phase, "kernel_ms_all": [round(float(x), 4) for x in kms]},
    }

    if not args.no_snapshot:
        out["load_s"] = snapshot_leg(eng, words, flags, res, load_triples_s)

    if args.compare_dedup or args.workload == "dense_conf":
        # body-pair dedup strategies on the same batch (eager passes, best of 3)
        cmp = {}
        eng.set_tuning("use_graph", 0)
        for mode, label in ((1, "bitmap_hbm_rows"), (3, "bitmap_lds"), (2, "sort_unique")):
            eng.set_tuning("dedup", mode)
            if mode == 3 and eng.get_tuning("pairs_dedup") != 2:
                continue   # the bitmap does not fit a block's LDS
            eng.evaluate(words, flags)
            times = []
            for _ in range(3):
                t = time.perf_counter()
                r = eng.evaluate(words, flags)
                times.append(time.perf_counter() - t)
            k = eng.kernel_ms()
            ds = eng.dedup_stats()
            same = bool(np.array_equal(r["bodyPairs"], res["bodyPairs"]) and np.array_equal(r["support"], res["support"]))
            cmp[label] = {"ms": round(min(times) * 1e3, 3), "pairs_wave_ms": round(float(k[3]), 3),
                          "sort_ms": round(ds["sort_ms"], 3), "keys_emitted": int(ds["keys_emitted"]),
                          "same_counts": same}
            log("dedup %s: %s" % (label, cmp[label]))
        eng.set_tuning("dedup", 0)
        eng.set_tuning("use_graph", 1)
        out["dedup_compare"] = cmp

    if not args.no_per_rule:
        t = time.perf_counter()
        pr = eng.evaluate(words, flags | L.PER_RULE)
        pr_wall = time.perf_counter() - t
        lat = pr["latencyUs"].astype(np.float64) / 1000.0
        same = bool(np.array_equal(pr["support"], res["support"]) and np.array_equal(pr["bodyPairs"], res["bodyPairs"]))
        out["per_rule"] = {"rules_per_s": round(n / pr_wall, 1), "wall_s": round(pr_wall, 3),
                           "p50_ms": round(pct(lat, 0.5), 4), "p99_ms": round(pct(lat, 0.99), 4),
                           "same_counts_as_batched": same}
        out["speedup_vs_per_rule"] = round(rules_per_s / (n / pr_wall), 2)
        out["latency_ms"]["per_rule_p50"] = out["per_rule"]["p50_ms"]
        out["latency_ms"]["per_rule_p99"] = out["per_rule"]["p99_ms"]
        log("per-rule launch: %.1f rules/s, p50 %.3f ms p99 %.3f ms, same counts: %s" % (
            n / pr_wall, out["per_rule"]["p50_ms"], out["per_rule"]["p99_ms"], same))

    if not args.no_cpu:
        t = time.time()
        out["cpu_baseline"] = cpu_baseline(wl, rules, args.cpu_sample or sample)
        log("cpu baseline %.1fs: %s" % (time.time() - t, out["cpu_baseline"]["value"]))

    batch.close()
    eng.close()
    print(json.dumps(out), flush=True)


def serve_mode(args):
    """Serving: a Poisson stream of rule requests (1-8 atoms, predicates
    Zipf s=1.0) against the resident high-fan-out graph, through the
    library's continuous-batching server; reports rules/s and p50/p99 latency
    from arrival to result."""
    from masterthesis_gpu_acceleration_amd import serving as S
    t0 = time.time()
    wl = W.hub_fanout()
    g = wl.graph
    arrivals, reqs = W.serving_requests(graph=g, n_requests=args.requests, rate_per_s=args.rate,
                                        max_work=args.max_work or None)
    kept = getattr(W.serving_requests, "last_kept", {})
    log("serving workload: %s; %d requests at %.0f/s, kept fraction per path length %s (%.1fs)" % (
        describe(wl), len(reqs), args.rate, kept, time.time() - t0))
    words = L.rule_words_for_graph(reqs, g)
    eng = L.Engine.from_graph(g)
    flags = 0     # reference semantics: body size only for rules with support > 0
    eng.evaluate(words[:256], flags)     # warm the kernels and the dedup rows
    srv = S.Server(eng, max_batch=2048, max_in_flight=3, flags=flags)
    try:
        r = S.run_stream(srv, words, arrivals,
                         progress=lambda sub, done: log("serving: %d submitted, %d completed" % (sub, done)))
        st = srv.stats()
    finally:
        srv.stop()
    lat = r["latency_ms"]
    comp = r["completions"]
    out = {
        "metric": METRIC, "value": round(r["rules_per_s"], 1), "unit": "rules/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(r["span_s"] * 1000.0, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "serving: Poisson stream of %d rule requests (1-8 atoms, predicates Zipf s=1.0, "
                               "paths with estimated search work above %.0e redrawn) at %.0f/s over %s; support for "
                               "every request, body pairs when support > 0"
                               % (len(reqs), args.max_work, args.rate, describe(wl)),
                   "requests": len(reqs), "rate_per_s": args.rate, "max_work": args.max_work,
                   "kept_fraction_by_path_length": {str(k): round(v, 4) for k, v in kept.items()}},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                       "max": round(float(lat.max()), 4)},
        "serving": st,
        "mean_batch": round(float(comp["batchSize"].mean()), 1),
        "rules_with_support": int((comp["result"]["support"] > 0).sum()),
    }
    log("serving: %.0f rules/s, p50 %.3f ms, p99 %.3f ms, %d batches" % (
        r["rules_per_s"], out["latency_ms"]["p50"], out["latency_ms"]["p99"], st["batches"]))
    # rate sweep: the first requests of the same stream, arrival times rescaled
    # to each rate (a Poisson stream stays Poisson under rescaling)
    sweep = []
    n = min(len(reqs), args.sweep_requests)
    for rate in [float(x) for x in args.rates.split(",") if x.strip()]:
        arr = np.asarray(arrivals[:n], np.float64) * (args.rate / rate)
        srv = S.Server(eng, max_batch=2048, max_in_flight=3, flags=flags)
        try:
            rr = S.run_stream(srv, words[:n], arr)
            ss = srv.stats()
        finally:
            srv.stop()
        lt = rr["latency_ms"]
        row = {"rate_per_s": rate, "requests": n, "rules_per_s": round(rr["rules_per_s"], 1),
               "p50_ms": round(float(np.percentile(lt, 50)), 3), "p99_ms": round(float(np.percentile(lt, 99)), 3),
               "max_ms": round(float(lt.max()), 3), "batches": ss["batches"],
               "mean_batch": round(float(rr["completions"]["batchSize"].mean()), 1)}
        sweep.append(row)
        log("serving sweep: %(rate_per_s).0f/s offered -> %(rules_per_s).0f rules/s, p50 %(p50_ms).1f ms, "
            "p99 %(p99_ms).1f ms, mean batch %(mean_batch).1f" % row)
    if sweep:
        out["rate_sweep"] = sweep
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

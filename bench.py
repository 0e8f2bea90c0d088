#!/usr/bin/env python3
"""Benchmark of the rule support / body-pair counting path on one MI355X.

Default run (`python bench.py --gpus 1 --steps K --warmup W`), one JSON line:

* headline -- DBpedia-shaped configuration: 206,000 triples over 50,000
  entities and 300 predicates (predicate frequency Zipf s=1.0, per-predicate
  out-degree lognormal(0,1)), 10,000 closed rules of 1-3 body atoms. A step is
  one evaluation of all rules (exact support and distinct body-pair count of
  every rule, also where support is 0 -- the work CPU_01_nonparallel does)
  from a batch compiled once and kept resident in HBM, replayed as a HIP
  graph. `value` = rules/s over K timed steps.
* "hub_fanout" -- the same measurement on the high-fan-out configuration
  (6.8M triples, 500,000 entities, 60 predicates, ~50 hubs of degree ~1e5,
  50 rules of 2-4 atoms), with its own ms_per_step, roofline and CPU baseline.
* "serving" -- the continuous-batching server over the resident high-fan-out
  graph: its capacity (the stream offered far above what it serves), the
  serving stream (1-8 atoms, predicates Zipf s=1.0, Poisson arrivals; no
  work filter) and, beside it, the same stream with paths of estimated search
  work above 1e8 redrawn; p50/p99/max latency from arrival to result, served
  rate, mean batch, requests routed to heavy batches.
* "mining" -- a level-wise rule miner (masterthesis_gpu_acceleration_amd.mining)
  over the headline graph: every candidate of one body length in one
  support-pruned device batch (gpu_evaluate_rules_min: body pairs only for
  candidates that meet minSupport and minHeadCoverage), per level
  candidates, survivors and times; beside it the same loop with body pairs
  counted for every closed candidate with support > 0 (thresholds applied
  on the host), which must find the same rules.

Every block also reports:
  * per_rule: the same rules through the reference library's per-rule launch
    sequence (support launch + sync + copy, body-size launch + sync + copy),
    with the reference bench's statistics of the per-rule device times
    (avg/median/std/min/max/P95, overall and by body size) and p99;
  * parity_vs_golden: the timed counts against the reference programs'
    outputs for the same seeded graph and rules (tests/golden/*.json);
  * roofline: algorithmic bytes of the dominant phase (host traffic model of
    the same plans) over the event-timed duration of the kernels that run
    that phase (thread-mode kernel + wave-mode kernel with its sub-item
    pass), against 8 TB/s; counters (traffic, L2 hit) and the traced phase
    time from the committed rocprofv3 summary (profiles/pmc_traffic.json);
  * tuning: each configuration runs at the tuning its committed profile
    was taken at (the autotuner's choice on MI355X), so the live figures and
    the profile describe the same launches; --autotune re-runs the tuner;
  * cpu_baseline: the reference's CPU_01_nonparallel program (built from the
    reference sources into oracle/_ref/, single-threaded), sharded over the
    host cores -- the rate on all cores and on one core.

The data is synthetic (the reference's datasets are not available), generated
from fixed seeds.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
       python bench.py --workload hub_fanout --only      (one configuration)
       python bench.py --serve                           (serving block only)
       python bench.py --mining                          (mining block only)
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Hardware queues per process (read when the HIP runtime starts): the
# library's two streams plus the server's three slot streams need five; with
# HIP's default of four, two slots share a queue and each waits for the
# other's batches.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

from masterthesis_gpu_acceleration_amd import lib as L  # noqa: E402
from masterthesis_gpu_acceleration_amd import workloads as W  # noqa: E402

METRIC = "rules/s & body-pairs/s (batched vs per-rule launch); p50/p99 rule latency"
HBM_PEAK_GBPS = 8000.0
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
CPU01 = os.path.join(ROOT, "oracle", "_ref", "cpu01")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(x, q):
    """The reference bench's percentile: sorted[ceil(q n) - 1]
    (GPU_03/bench_gpu.cu:725-729)."""
    x = np.sort(np.asarray(x, np.float64))
    if x.size == 0:
        return 0.0
    k = max(0, min(x.size - 1, int(np.ceil(q * x.size)) - 1))
    return float(x[k])


def stats(x):
    """avg/median/std/min/max/P95 as GPU_03/bench_gpu.cu:714-760 prints them, plus p99."""
    x = np.asarray(x, np.float64)
    if x.size == 0:
        return {}
    return {"n": int(x.size), "avg": round(float(x.mean()), 5), "median": round(float(np.median(x)), 5),
            "std": round(float(x.std()), 5), "min": round(float(x.min()), 5), "max": round(float(x.max()), 5),
            "p95": round(pct(x, 0.95), 5), "p99": round(pct(x, 0.99), 5)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may use (the box's CPU share)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


# name: (generator, golden fixture, CPU-baseline sample in rules (0: all))
WORKLOADS = {
    "dbpedia_shaped": (lambda: W.dbpedia_shaped(), "dbpedia_shaped_full", 0),
    "hub_fanout": (lambda: W.hub_fanout(), "hub_fanout_full", 16),
    "small_paths": (lambda: W.small_paths(), "small_paths_full", 0),
    "dense_conf": (lambda: W.dense_conf(), "dense_conf_full", 0),
}


def describe(wl):
    g = wl.graph
    nb = [len(r[1]) for r in wl.rules]
    return "%s: %d triples, %d entities, %d predicates, %d rules of %d-%d atoms" % (
        wl.name, g.n_triples, g.n_entities, g.n_predicates, len(wl.rules), min(nb), max(nb))


# ------------------------------------------------------------ parity

def graph_sha(g):
    h = hashlib.sha256()
    for a in (g.s, g.p, g.o):
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


def head_pair_in_body(rule):
    head, body = rule
    bv = {t[1] for a in body for t in (a[0], a[2]) if t[0] == "v"}
    return all(t[0] == "v" and t[1] in bv for t in (head[0], head[2]))


def parity_vs_golden(name, wl, res, gated):
    """The counts of the timed batch against the reference programs' outputs
    for the same graph and rules (CPU_01_nonparallel's support, head figures
    and body size; body pairs are 0 where a head variable is absent from the
    body, as Rewritten_CPP counts them)."""
    path = os.path.join(GOLDEN_DIR, name + ".json")
    try:
        fx = json.load(open(path))
    except OSError:
        return {"fixture": name, "ok": None, "note": "fixture absent"}
    if fx["graph_sha256"] != graph_sha(wl.graph) or fx["rules_text"] != [W.rule_text(r) for r in wl.rules]:
        return {"fixture": name, "ok": False, "note": "generated workload differs from the fixture's"}
    bad = 0
    for i, r in enumerate(wl.rules):
        sup, hsize, hsup, body = fx["cpu01"][i]
        pairs = body if head_pair_in_body(r) else 0
        got = res[i]
        want = [sup, hsize, hsup, pairs if sup > 0 else 0]
        have = [int(got["support"]), int(got["headSize"]), int(got["headSupport"]), int(got["bodySize"])]
        if not gated:
            want.append(pairs)
            have.append(int(got["bodyPairs"]))
        bad += int(want != have or int(got["status"]) != 0)
    return {"fixture": name, "rules": len(wl.rules), "mismatches": bad, "ok": bad == 0,
            "compared": "support, headSize, headSupport, bodySize" + ("" if gated else ", bodyPairs")
                        + " vs CPU_01_nonparallel"}


# ------------------------------------------------------------ CPU baseline

def cpu_baseline(wl, sample):
    """The reference CPU_01_nonparallel program (single-threaded), one
    process per host core, each on a share of the rules (round robin) of the
    first `sample` rules (0: all). Rule compute = sum of its per-rule support
    + confidence times, as the reference notebooks count it (thesis p.43):
    the one-core rate is rules / total rule compute, the all-core rate rules
    / the largest process's rule compute. Falls back to the project's CPU
    oracle on one core when the reference build is absent."""
    rules = wl.rules if sample <= 0 else wl.rules[:sample]
    cores = max(1, min(host_cores(), 16, len(rules)))
    if not os.path.exists(CPU01):
        from oracle import Oracle  # the port, when the reference build is absent
        g = wl.graph
        o = Oracle(g.s, g.p, g.o)
        t = time.time()
        o.eval(W.encode_rules(rules))
        dt = time.time() - t
        return {"value": round(len(rules) / dt, 2), "unit": "rules/s", "cores": 1, "kind": "port",
                "sample": "%d of the %d rules, full graph; project CPU oracle on one core; host %s" % (
                    len(rules), len(wl.rules), cpu_model())}
    with tempfile.TemporaryDirectory() as d:
        ttl = os.path.join(d, "g.ttl")
        W.write_ttl(wl.graph, ttl)
        shards = [rules[k::cores] for k in range(cores)]

        def run(k):
            rf = os.path.join(d, "r%02d.txt" % k)
            W.write_rules(shards[k], rf)
            t = time.time()
            out = subprocess.run([CPU01, ttl, rf, "--json"], capture_output=True, text=True, timeout=400, check=True)
            return json.loads(out.stdout), time.time() - t

        t = time.time()
        with ThreadPoolExecutor(cores) as ex:
            runs = list(ex.map(run, range(cores)))
        wall = time.time() - t
    per_proc = [sum(r["supportMs"] + r["confidenceMs"] for r in rows) for rows, _ in runs]
    total_ms = sum(per_proc)
    n = sum(len(rows) for rows, _ in runs)
    return {"value": round(n / (max(per_proc) / 1000.0), 2), "unit": "rules/s", "cores": cores, "kind": "reference",
            "single_core": round(n / (total_ms / 1000.0), 2),
            "sample": "%s of the %d rules, full graph; CPU_01_nonparallel, %d single-threaded processes (one per core, "
                      "rules round robin); rule compute %.1f ms in all, %.1f ms in the largest share; wall %.1f s "
                      "incl. each process's Turtle load; host %s, %d cores available" % (
                          "all" if n == len(wl.rules) else "first %d" % n, len(wl.rules), cores, total_ms,
                          max(per_proc), wall, cpu_model(), host_cores())}


# ------------------------------------------------------------ one configuration

PROFILE_TABLE = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def profile_entry(workload):
    """The committed profile of a configuration (rocprofv3 kernel trace and
    counter runs, summarised into profiles/pmc_traffic.json), or None."""
    try:
        d = json.load(open(PROFILE_TABLE))[workload]
        d["phase_hbm_bytes"], d["tuning"]
    except (OSError, KeyError, ValueError, TypeError):
        return None
    return d


def pmc_traffic(workload, tuning):
    """HBM bytes per pass of the dominant phase's kernels from the committed
    counter summary (FETCH_SIZE and WRITE_SIZE runs, scaled by the copy
    yardstick), the L2 hit rate, the phase's kernel time in the profiled
    run, and whether the profile was taken at this tuning."""
    d = profile_entry(workload)
    if d is None:
        return None
    prof_t = d["tuning"]
    same = all(prof_t.get(k) == tuning.get(k) for k in TUNE_SET)
    return {"traffic": int(d["phase_hbm_bytes"]), "l2_hit": d["l2_hit"], "traffic_source": d["source"],
            "profile_phase_ms": d.get("phase_ms"), "profile_frac": d.get("frac"),
            "profile_kernels": d.get("phase_kernels"), "traffic_tuning": prof_t, "traffic_tuning_matches": same}


# settable tuning keys, in the order --tune takes them
TUNE_SET = ("budget", "thread_waves_per_cu", "wave_waves_per_cu", "dedup", "giant_after", "chunks_per_wave", "wave_engine")


def tune_engine(eng, name, words, flags, args):
    """Fixed tuning (--tune), the autotuner (--autotune, or when no profile of
    this configuration is committed), else the tuning the committed profile
    was taken at (the autotuner's choice on MI355X). Returns the source."""
    prof = profile_entry(name)
    if args.tune:
        knobs = [int(x) for x in args.tune.split(",")]
        for key, val in zip(TUNE_SET, knobs):
            eng.set_tuning(key, val)
        log("fixed tuning: %s" % args.tune)
        return "--tune " + args.tune
    if args.no_autotune:
        return "library defaults"
    if prof is not None and not args.autotune:
        for key in TUNE_SET:
            if key in prof["tuning"]:
                eng.set_tuning(key, int(prof["tuning"][key]))
        log("tuning of the committed profile: %s" % prof["tuning"])
        return "profiles/pmc_traffic.json (the autotuner's choice when the profile was taken)"
    t = time.time()
    tl = eng.autotune(words, flags)
    best = tl[np.argmin(tl[:, 3])]
    log("autotune (%.1fs, %d trials): budget %d, thread waves/CU %d, wave waves/CU %d -> %.3f ms; chunks/wave %d, "
        "giant_after %d, dedup %d" % (time.time() - t, len(tl), best[0], best[1], best[2], best[3],
                                      eng.get_tuning("chunks_per_wave"), eng.get_tuning("giant_after"),
                                      eng.get_tuning("dedup")))
    return "autotuner (%d trials, %.1f s)" % (len(tl), time.time() - t)


def time_batch(eng, words, flags, steps, warmup):
    """ms per step of K graph replays of a resident batch (each step waits for
    its results), device ms per replay, and the last results."""
    batch = eng.batch(words, flags)
    try:
        for _ in range(warmup):
            res = batch.run()
        dev = []
        t = time.perf_counter()
        for _ in range(steps):
            res = batch.run()
            dev.append(eng.last_stats()["total_ms"])
        wall = time.perf_counter() - t
    finally:
        batch.close()
    return wall * 1000.0 / steps, dev, res.copy()


def measure(args, name, wl, decompose):
    """One configuration's block of the JSON line."""
    gen, golden, sample = WORKLOADS[name]
    g = wl.graph
    rules = wl.rules
    n = len(rules)
    words = L.rule_words_for_graph(rules, g)
    t = time.time()
    eng = L.Engine.from_graph(g)
    load_s = time.time() - t
    log("%s: graph resident in HBM (%.2fs)" % (name, load_s))
    flags = 0 if args.gated else L.UNGATED
    try:
        if args.no_overlap:
            eng.set_tuning("overlap", 0)
        if args.no_fuse:
            eng.set_tuning("fuse_support", 0)
        for kv in args.set:
            key, val = kv.split("=", 1)
            eng.set_tuning(key, int(val))
        tuning_source = tune_engine(eng, name + ("_gated" if args.gated else ""), words, flags, args)
        tuning = {k: eng.get_tuning(k) for k in TUNE_SET}
        t = time.time()
        ms_step, dev_ms, res = time_batch(eng, words, flags, args.steps, args.warmup)
        log("%s: timed %d steps: %.3f ms/step (device %.3f ms)" % (name, args.steps, ms_step, float(np.mean(dev_ms))))
        rules_per_s = n / (ms_step / 1000.0)
        pairs = int(res["bodyPairs"][res["bodyPairs"] > 0].sum())
        # the library's counter keys, and what they count under each wave engine
        keys = ("seeds", "steps", "donations", "leaf_draws", "draws", "busy_lane_steps", "draw_lanes")
        labels = keys if eng.get_tuning("wave_engine") == 0 else (
            "seeds", "steps", "fill_steps", "streamed_steps", "balanced_steps", "candidate_slots", "candidates_passed")
        wave = {ph: {lab: eng.get_tuning("wave_%s_%s" % (ph, q)) for q, lab in zip(keys, labels)}
                for ph in ("support", "pairs")}
        cyc = {q: eng.get_tuning("wave_pairs_cycles_" + q) for q in ("take", "setup", "leaf_draw", "draw", "own",
                                                                     "step_end", "seed_end", "sub_items")}
        if any(cyc.values()):   # the profiling build (GPU_RULES_LIB=.../libgpu_rules_prof.so)
            wave["pairs_cycles"] = cyc
        wave["giants"] = eng.get_tuning("last_giants")
        searches = {"support": eng.get_tuning("last_support_searches"), "pairs": eng.get_tuning("last_pair_searches"),
                    "support_fused_rules": eng.get_tuning("last_fused"),
                    "rules_sharing_a_search": eng.get_tuning("last_shared")}
        block = {
            "value": round(rules_per_s, 1), "unit": "rules/s", "ms_per_step": round(ms_step, 4),
            "config": {"workload": describe(wl) + (", exact support + distinct body pairs for %s, resident batch, "
                                                    "HIP graph replay" % ("rules with support > 0" if args.gated
                                                                          else "every rule")),
                       "rules": n, "triples": g.n_triples},
            "body_pairs_per_s": round(pairs / (ms_step / 1000.0), 1), "body_pairs_total": pairs,
            "rules_with_support": int((res["support"] > 0).sum()),
            "escalated_seeds": int(eng.last_stats()["heavy_seeds"]),
            "searches": searches,
            "wave_mode": wave,
            "tuning": dict(tuning, pairs_dedup=eng.get_tuning("pairs_dedup"), wave_slots=eng.get_tuning("wave_slots"),
                          fuse_active=eng.get_tuning("fuse_active"), source=tuning_source),
            "latency_ms": {"batched_p50": round(pct(dev_ms, 0.5), 4), "batched_p99": round(pct(dev_ms, 0.99), 4)},
            "parity_vs_golden": parity_vs_golden(golden, wl, res, args.gated),
        }
        log("%s: parity vs golden: %s" % (name, block["parity_vs_golden"]))

        # pre-dedup groundings: what the reference's body-size kernel would emit for these rules,
        # counted on the host (this engine never enumerates them); too many to count for high fan-out
        if name != "hub_fanout":
            t = time.perf_counter()
            gr = eng.body_groundings(words)
            cnt = int(gr[gr >= 0].sum())
            block["body_groundings"] = {
                "per_step": cnt, "rules_counted": int((gr >= 0).sum()),
                "per_s_equivalent": round(cnt / (ms_step / 1000.0), 1),
                "what": "injective body groundings of the rules (one pair each in the reference's body-size kernel, "
                        "before dedup), counted on the host; per second of this engine's step",
                "count_s": round(time.perf_counter() - t, 3)}
            log("%s: %d body groundings per step (%.3g per second of the step)" % (name, cnt, cnt / (ms_step / 1000.0)))

        # end to end: compile + upload + run + readback of a fresh batch (PCIe-inclusive; never `value`)
        t = time.perf_counter()
        eng.evaluate(words, flags)
        block["e2e_rules_per_s"] = round(n / (time.perf_counter() - t), 1)

        # roofline of the dominant phase: eager passes with HIP events on the library's stream
        eng.set_tuning("use_graph", 0)
        kms = np.zeros(4)
        reps = 3
        for _ in range(reps):
            eng.evaluate(words, flags)
            kms += eng.kernel_ms()
        kms /= reps
        eng.set_tuning("use_graph", 1)
        traffic = eng.traffic(words)
        phase_ms = [kms[0] + kms[1], kms[2] + kms[3]]
        phase = int(np.argmax(phase_ms))
        bytes_phase = int(traffic[:, phase].sum())
        achieved = bytes_phase / (phase_ms[phase] * 1e-3) / 1e9 if phase_ms[phase] > 0 else 0.0
        kernels = ["k_thread_batch<%d>" % phase, "k_wave_batch<%d>" % phase, "k_giant_expand", "k_wave_sub<%d>" % phase]
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": None,
                "phase": ("support", "body pairs")[phase], "kernels": kernels,
                "phase_ms": round(float(phase_ms[phase]), 4), "algorithmic_bytes": bytes_phase,
                "kernel_ms": {"thread": round(float(kms[2 * phase]), 4),
                              "wave_and_sub_items": round(float(kms[2 * phase + 1]), 4)},
                "kernel_ms_all": [round(float(x), 4) for x in kms]}
        pm = pmc_traffic(name + ("_gated" if args.gated else ""), tuning)
        if pm:
            roof.update(pm)
            if pm.get("profile_phase_ms"):
                # HBM bytes the counters saw over the same kernels' traced time
                roof["counter_GBps"] = round(pm["traffic"] / (pm["profile_phase_ms"] * 1e-3) / 1e9, 2)
                roof["phase_ms_vs_profile"] = round(float(phase_ms[phase]) / pm["profile_phase_ms"], 4)
        block["roofline"] = roof
        log("%s: kernels ms %s; phase %s %.3f ms, %.1f GB/s model bytes" % (
            name, [round(float(x), 3) for x in kms], roof["phase"], phase_ms[phase], achieved))

        if decompose:
            # what batching gains without sharing equal bodies' searches and without counting support in the
            # body-pair pass (same rules, same tuning)
            dec = {}
            for label, knobs in (("no_shared_bodies", {"share_bodies": 0}), ("no_fused_support", {"fuse_support": 0}),
                                 ("neither", {"share_bodies": 0, "fuse_support": 0})):
                for k, v in knobs.items():
                    eng.set_tuning(k, v)
                ms, _, r2 = time_batch(eng, words, flags, args.steps, 2)
                dec[label] = {"rules_per_s": round(n / (ms / 1000.0), 1), "ms_per_step": round(ms, 4),
                              "pair_searches": eng.get_tuning("last_pair_searches"),
                              "support_searches": eng.get_tuning("last_support_searches"),
                              "same_counts": bool(np.array_equal(r2["support"], res["support"]) and
                                                  np.array_equal(r2["bodyPairs"], res["bodyPairs"]))}
                for k in knobs:
                    eng.set_tuning(k, 1)
                log("%s: %s: %s" % (name, label, dec[label]))
            block["decomposition"] = dec

        if not args.no_snapshot:
            block["load_s"] = snapshot_leg(eng, words, flags, res, load_s)

        if args.compare_dedup or name == "dense_conf":
            block["dedup_compare"] = dedup_compare(eng, words, flags, res)

        if not args.no_per_rule:
            t = time.perf_counter()
            pr = eng.evaluate(words, flags | L.PER_RULE)
            pr_wall = time.perf_counter() - t
            sup_ms, conf_ms = eng.rule_times(n)
            lat = pr["latencyUs"].astype(np.float64) / 1000.0
            nb = np.array([len(r[1]) for r in rules])
            same = bool(np.array_equal(pr["support"], res["support"]) and
                        np.array_equal(pr["bodyPairs"], res["bodyPairs"]))
            block["per_rule"] = {
                "rules_per_s": round(n / pr_wall, 1), "wall_s": round(pr_wall, 3),
                "p50_ms": round(pct(lat, 0.5), 4), "p99_ms": round(pct(lat, 0.99), 4),
                "same_counts_as_batched": same,
                "host_wall_ms": stats(lat),
                "support_ms": stats(sup_ms), "confidence_ms": stats(conf_ms),
                "by_body_size": {str(k): {"support_ms": stats(sup_ms[nb == k]), "confidence_ms": stats(conf_ms[nb == k])}
                                 for k in sorted(set(nb.tolist()))}}
            block["speedup_vs_per_rule"] = round(rules_per_s / (n / pr_wall), 2)
            block["latency_ms"]["per_rule_p50"] = block["per_rule"]["p50_ms"]
            block["latency_ms"]["per_rule_p99"] = block["per_rule"]["p99_ms"]
            log("%s: per-rule launch: %.1f rules/s, p50 %.3f ms p99 %.3f ms, same counts: %s" % (
                name, n / pr_wall, block["per_rule"]["p50_ms"], block["per_rule"]["p99_ms"], same))
    finally:
        eng.close()

    if not args.no_cpu:
        t = time.time()
        block["cpu_baseline"] = cpu_baseline(wl, args.cpu_sample or sample)
        log("%s: cpu baseline %.1fs: %s rules/s on %d cores" % (name, time.time() - t, block["cpu_baseline"]["value"],
                                                               block["cpu_baseline"]["cores"]))
    return block


def snapshot_leg(eng, words, flags, res, load_triples_s):
    """Graph load times: from term-id triples (host or device sort + index
    build + upload) against a binary snapshot of the same graph (map + checks
    + upload); the snapshot's context must give the same counts."""
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
    return {"from_triples": round(load_triples_s, 3), "snapshot_save": round(save_s, 3),
            "from_snapshot": round(snap_s, 3), "snapshot_mb": round(size / 1e6, 1), "same_counts": same}


def dedup_compare(eng, words, flags, res):
    """Body-pair dedup strategies on the same batch (eager passes, best of 3)."""
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
                      "sort_ms": round(ds["sort_ms"], 3), "keys_emitted": int(ds["keys_emitted"]), "same_counts": same}
        log("dedup %s: %s" % (label, cmp[label]))
    eng.set_tuning("dedup", 0)
    eng.set_tuning("use_graph", 1)
    return cmp


# ------------------------------------------------------------ serving

def serve_leg(eng, g, n, rate, max_work, label):
    """The first n requests of the 100,000-request serving stream (1-8 atoms,
    predicates Zipf s=1.0), arrivals Poisson at `rate`, through the server;
    max_work: paths of larger estimated search work are redrawn (None: kept)."""
    from masterthesis_gpu_acceleration_amd import serving as S
    t = time.time()
    arrivals, reqs = W.serving_requests(graph=g, n_requests=100000, rate_per_s=rate, max_work=max_work)
    kept = getattr(W.serving_requests, "last_kept", {}) if max_work else {}
    arrivals, reqs = arrivals[:n], reqs[:n]
    words = L.rule_words_for_graph(reqs, g)
    log("serving %s: %d requests at %.0f/s (%.1fs to generate)" % (label, n, rate, time.time() - t))
    srv = S.Server(eng, max_batch=2048, max_in_flight=3, flags=0)
    try:
        r = S.run_stream(srv, words, arrivals, progress=lambda sub, done: log(
            "serving %s: %d submitted, %d completed" % (label, sub, done)))
        st = srv.stats()
    finally:
        srv.stop()
    lat = r["latency_ms"]
    comp = r["completions"]
    nb = np.array([len(q[1]) for q in reqs])
    out = {"requests": n, "offered_per_s": rate, "served_rules_per_s": round(r["rules_per_s"], 1),
           "heavy_requests": st.get("heavy_requests"), "heavy_batches": st.get("heavy_batches"),
           "span_s": round(r["span_s"], 3),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                          "max": round(float(lat.max()), 3)},
           "latency_p99_ms_by_atoms": {str(k): round(float(np.percentile(lat[nb == k], 99)), 3)
                                       for k in sorted(set(nb.tolist()))},
           "mean_batch": round(float(comp["batchSize"].mean()), 1), "batches": st["batches"],
           "refused": st["refused"], "device_failure": st["device_failure"],
           "errors": int((comp["result"]["status"] == 2).sum()),
           "rules_with_support": int((comp["result"]["support"] > 0).sum()),
           "work_filter": ("paths of estimated search work above %.0e redrawn; kept fraction per path length %s" % (
               max_work, {str(k): round(v, 4) for k, v in kept.items()})) if max_work else "none"}
    log("serving %s: %.0f rules/s served, p50 %.1f ms, p99 %.1f ms, max %.1f ms, mean batch %.1f" % (
        label, r["rules_per_s"], out["latency_ms"]["p50"], out["latency_ms"]["p99"], out["latency_ms"]["max"],
        out["mean_batch"]))
    return out


def serving_block(args, g):
    eng = L.Engine.from_graph(g)
    try:
        if args.serve_heavy_work >= 0:
            eng.set_tuning("serve_heavy_work", int(args.serve_heavy_work))
        eng.evaluate(L.rule_words_for_graph(W.serving_stream(256, graph=g).rules, g), 0)   # warm the kernels
        blk = {"config": "continuous-batching server (3 slots, batches up to 2048 requests) over the resident "
                         "high-fan-out graph (6.8M triples); support for every request, body pairs when support > 0 "
                         "(reference semantics); latency from scheduled arrival to result on the host; requests of "
                         "estimated search work above serve_heavy_work in heavy batches of their own (one in flight, "
                         "serve_heavy_percent of the grid)",
               "serve_heavy_work": eng.get_tuning("serve_heavy_work"),
               "serve_heavy_percent": eng.get_tuning("serve_heavy_percent")}
        if args.serve_capacity_requests > 0:
            # offered far above what the device serves: the served rate is the capacity
            blk["capacity"] = serve_leg(eng, g, args.serve_capacity_requests, args.serve_capacity_rate, None,
                                        "capacity")
        blk["stream"] = serve_leg(eng, g, args.serve_requests, args.serve_rate, None, "stream")
        if args.serve_filtered_requests > 0:
            blk["stream_work_filtered"] = serve_leg(eng, g, args.serve_filtered_requests, args.serve_filtered_rate, 1e8,
                                                    "filtered")
    finally:
        eng.close()
    return blk


# ------------------------------------------------------------ main

# ------------------------------------------------------------ mining

class _UnprunedEvaluator:
    """Body pairs for every closed candidate with support > 0 (a gated
    batch); open candidates for support only -- a miner that applies its
    thresholds on the host."""

    def __init__(self, eng, g):
        self.eng, self.g = eng, g
        self.device_ms = []

    def __call__(self, records):
        from masterthesis_gpu_acceleration_amd import mining as M
        words = M.words_of_keys(records, self.g.n_entities)
        closed = M.closed_mask(records)
        out = np.zeros(len(records), L.RESULT64_DTYPE)
        ms = 0.0
        if closed.any():
            out[closed] = self.eng.evaluate(words[closed], 0)
            ms += self.eng.last_stats()["total_ms"]
        if (~closed).any():
            out[~closed] = self.eng.evaluate_min(words[~closed], M.SUPPORT_ONLY, 0.0)
            ms += self.eng.last_stats()["total_ms"]
        self.device_ms.append(ms)
        return out


def mining_block(args, g):
    from masterthesis_gpu_acceleration_amd import mining as M
    heads = M.frequent_predicates(g, args.mine_heads)
    bodies = M.frequent_predicates(g, args.mine_bodies)
    miner = M.Miner(heads, bodies, args.mine_atoms, args.mine_support, args.mine_hc)
    cfg = {"graph": "%d triples, %d entities, %d predicates" % (len(g.s), g.n_entities, g.n_predicates),
           "head_predicates": len(heads), "body_predicates": len(bodies), "max_body_atoms": args.mine_atoms,
           "min_support": args.mine_support, "min_head_coverage": args.mine_hc,
           "predicates": "the most frequent ones", "rules": "variables only, closed rules reported"}
    eng = L.Engine.from_graph(g)
    try:
        ev = M.GpuEvaluator(eng, g, miner)
        mined, st = miner.mine(ev)
        ev0 = _UnprunedEvaluator(eng, g)
        mined0, st0 = miner.mine(ev0)
    finally:
        eng.close()
    same = M.results_table(mined) == M.results_table(mined0) and st.candidates == st0.candidates
    levels = [{"body_atoms": k + 1, "candidates": c, "passed": p, "eval_ms": round(1e3 * e, 2),
               "device_ms": round(d, 3)}
              for k, (c, p, e, d) in enumerate(zip(st.candidates, st.passed, st.eval_s, ev.device_ms))]
    ncand = sum(st.candidates)
    out = {"config": cfg, "levels": levels, "mined_rules": len(mined), "candidates": ncand,
           "total_s": round(st.total_s, 3), "eval_s": round(sum(st.eval_s), 3),
           "device_ms": round(sum(ev.device_ms), 3), "host_refinement_s": round(st.total_s - sum(st.eval_s), 3),
           "candidates_per_s_evaluated": round(ncand / max(sum(st.eval_s), 1e-9), 1),
           "without_device_thresholds": {"eval_s": round(sum(st0.eval_s), 3),
                                         "device_ms": round(sum(ev0.device_ms), 3), "same_rules": same}}
    log("mining: %d candidates over %d levels, %d rules; evaluation %.2f s (device %.1f ms), refinement %.2f s; "
        "without device thresholds: evaluation %.2f s (device %.1f ms), same rules: %s" % (
            ncand, len(levels), len(mined), sum(st.eval_s), sum(ev.device_ms), out["host_refinement_s"],
            sum(st0.eval_s), sum(ev0.device_ms), same))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dbpedia_shaped", choices=sorted(WORKLOADS))
    ap.add_argument("--only", action="store_true", help="measure --workload alone (no high-fan-out or serving block)")
    ap.add_argument("--atoms", default="", help="diagnostics: keep only rules with these body lengths (comma list)")
    ap.add_argument("--cpu-sample", type=int, default=0, help="rules in the CPU-baseline sample (0: per workload)")
    ap.add_argument("--gated", action="store_true", help="body pairs only for rules with support > 0")
    ap.add_argument("--serve", action="store_true", help="only the serving block")
    ap.add_argument("--no-serve", action="store_true", help="leave out the serving block")
    ap.add_argument("--mining", action="store_true", help="only the mining block")
    ap.add_argument("--no-mining", action="store_true", help="leave out the mining block")
    ap.add_argument("--mine-heads", type=int, default=10, help="mining: head predicates (most frequent)")
    ap.add_argument("--mine-bodies", type=int, default=20, help="mining: body predicates (most frequent)")
    ap.add_argument("--mine-atoms", type=int, default=3)
    ap.add_argument("--mine-support", type=int, default=10)
    ap.add_argument("--mine-hc", type=float, default=0.01)
    ap.add_argument("--serve-requests", type=int, default=9000)
    ap.add_argument("--serve-rate", type=float, default=150.0)
    ap.add_argument("--serve-capacity-requests", type=int, default=3000,
                    help="requests of the capacity leg (offered at --serve-capacity-rate; 0: no capacity leg)")
    ap.add_argument("--serve-capacity-rate", type=float, default=2000.0)
    ap.add_argument("--serve-heavy-work", type=float, default=-1,
                    help="serve_heavy_work tuning for the server (-1: the library default; 0: no heavy batches)")
    ap.add_argument("--serve-filtered-requests", type=int, default=30000)
    ap.add_argument("--serve-filtered-rate", type=float, default=1000.0)
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
    ap.add_argument("--no-decompose", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="library default tuning")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner even where a committed profile names the tuning it was taken at")
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

    t0 = time.time()
    if args.serve:
        g = W.hub_fanout().graph
        print(json.dumps({"metric": METRIC, "serving": serving_block(args, g)}), flush=True)
        return

    if args.mining:
        g = WORKLOADS[args.workload][0]().graph
        print(json.dumps({"metric": METRIC, "mining": mining_block(args, g)}), flush=True)
        return

    wl = WORKLOADS[args.workload][0]()
    if args.atoms:
        keep = {int(x) for x in args.atoms.split(",")}
        wl = W.Workload(wl.graph, [r for r in wl.rules if len(r[1]) in keep], wl.name)
    log("workload %s (%.1fs)" % (describe(wl), time.time() - t0))
    head = measure(args, args.workload, wl, decompose=not args.no_decompose)
    out = {
        "metric": METRIC, "value": head.pop("value"), "unit": "rules/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head.pop("ms_per_step"), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
    }
    out.update(head)
    if not args.only and not args.no_mining:
        out["mining"] = mining_block(args, wl.graph)
    if not args.only and args.workload != "hub_fanout":
        t = time.time()
        hub = W.hub_fanout()
        log("workload %s (%.1fs)" % (describe(hub), time.time() - t))
        out["hub_fanout"] = measure(args, "hub_fanout", hub, decompose=False)
        if not args.no_serve:
            out["serving"] = serving_block(args, hub.graph)
    log("bench done in %.1fs" % (time.time() - t0))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

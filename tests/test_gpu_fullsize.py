"""GPU parity at the full configuration sizes, against the reference programs'
own outputs (CPU_01_nonparallel and Rewritten_CPP, compiled from the
reference sources; tests/golden/make_golden.py):

  * high-fan-out graph (6.8M triples, 500k entities, 60 predicates, ~50 hubs
    of degree ~1e5) with its 50 rules of 2-4 atoms (hub_fanout_full);
  * dense confidence stress (100k triples, 2k entities, 32 rules of 3-8
    atoms) (dense_conf_full);
  * the serving stream over the high-fan-out graph: the first 400 requests
    of the 100,000-request stream, 1-8 atoms, no work filter
    (serving_stream_400), and the first 2,000 served alone, mixed and as one
    batch.

Each configuration runs with the tuning the autotuner picks, gated and
ungated, under every body-pair dedup mode (HBM bitmap rows, LDS bitmaps,
sort-unique), and with seeds cut into sub-items after 512 wave-mode steps.
Every count is an exact integer: tolerance 0.
"""
import sys
import time

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import serving as S
from masterthesis_gpu_acceleration_amd import workloads as W

import goldens
from test_gpu_parity import check

pytestmark = pytest.mark.gpu

FULL = ["hub_fanout_full", "dense_conf_full"]
FIELDS = ("support", "headSize", "headSupport", "bodySize", "bodyPairs", "status")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def engines():
    cache = {}
    yield cache
    for entry in cache.values():
        entry[0].close()


def engine_for(cache, name):
    if name not in cache:
        fx, g, rules = goldens.load(name)
        eng = L.Engine.from_graph(g)
        cache[name] = (eng, fx, g, rules, L.rule_words_for_graph(rules, g))
    return cache[name]


def reset(eng):
    for k, v in (("dedup", 0), ("budget", 4096), ("giant_after", 2048), ("use_graph", 1), ("thread_waves_per_cu", 16),
                 ("wave_waves_per_cu", 8), ("chunks_per_wave", 2), ("fuse_support", 1), ("share_bodies", 1)):
        eng.set_tuning(k, v)


@pytest.mark.parametrize("name", FULL)
def test_autotuned(engines, name):
    """The tuning the bench runs with: autotuned on the whole batch."""
    say("autotuned", name)
    eng, fx, g, rules, words = engine_for(engines, name)
    try:
        eng.autotune(words, L.UNGATED)
        say("  tuning", {k: eng.get_tuning(k) for k in ("budget", "thread_waves_per_cu", "wave_waves_per_cu",
                                                       "pairs_dedup", "giant_after", "chunks_per_wave")})
        check(fx, rules, eng.evaluate(words, L.UNGATED), False, name + "/autotuned")
        check(fx, rules, eng.evaluate(words, 0), True, name + "/autotuned-gated")
        b = eng.batch(words, L.UNGATED)
        try:
            for _ in range(3):   # graph replays
                check(fx, rules, b.run(), False, name + "/autotuned-replay")
        finally:
            b.close()
    finally:
        reset(eng)


@pytest.mark.parametrize("name", FULL)
@pytest.mark.parametrize("dedup", [1, 3])
def test_bitmap_dedup_modes(engines, name, dedup):
    """HBM bitmap rows (1) and LDS bitmaps (3), with and without support
    counted in the body-pair pass."""
    say("dedup", name, dedup)
    eng, fx, g, rules, words = engine_for(engines, name)
    try:
        eng.set_tuning("dedup", dedup)
        for fuse in (1, 0):
            eng.set_tuning("fuse_support", fuse)
            check(fx, rules, eng.evaluate(words, L.UNGATED), False, "%s/dedup%d/fuse%d" % (name, dedup, fuse))
        check(fx, rules, eng.evaluate(words, 0), True, "%s/dedup%d/gated" % (name, dedup))
    finally:
        reset(eng)


def sort_subset(fx, rules, limit):
    """Rules whose distinct body pairs sum to at most `limit` (the sort-unique
    path holds every wave-mode seed's keys of a batch at once, at most 2^31),
    fewest pairs first so the subset keeps as many rules as fit."""
    order = sorted(range(len(rules)), key=lambda i: fx["cpu01"][i][3])
    keep, tot = [], 0
    for i in order:
        b = fx["cpu01"][i][3]
        if tot + b <= limit:
            keep.append(i)
            tot += b
    return sorted(keep), tot


@pytest.mark.parametrize("name", FULL)
def test_sort_dedup(engines, name):
    """Sort-unique dedup (2). On the high-fan-out graph the whole batch has
    5.4e9 distinct pairs, beyond the key buffers' 2^31; it runs on as many
    rules as fit 4e8 distinct pairs (hub paths of 2-4 atoms among them)."""
    say("sort-dedup", name)
    eng, fx, g, rules, words = engine_for(engines, name)
    idx, tot = sort_subset(fx, rules, 400_000_000)
    say("  %d of %d rules, %d distinct pairs" % (len(idx), len(rules), tot))
    assert len(idx) >= min(len(rules), 20)
    sub_fx = dict(fx, cpu01=[fx["cpu01"][i] for i in idx], rules_text=[fx["rules_text"][i] for i in idx])
    sub_rules = [rules[i] for i in idx]
    try:
        eng.set_tuning("dedup", 2)
        assert eng.get_tuning("sort_active") == 1
        res = eng.evaluate(words[idx], L.UNGATED)
        check(sub_fx, sub_rules, res, False, name + "/sort")
        say("  keys emitted", eng.dedup_stats()["keys_emitted"])
        check(sub_fx, sub_rules, eng.evaluate(words[idx], 0), True, name + "/sort-gated")
    finally:
        reset(eng)


def test_sort_capacity_limit_fails_loudly(engines):
    """A pass that emits more keys than the largest allowed capacity fails
    (gpu_evaluate_rules returns -1) instead of returning truncated counts; the
    context stays usable."""
    say("sort-cap-max")
    eng, fx, g, rules, words = engine_for(engines, "dense_conf_full")
    try:
        eng.set_tuning("dedup", 2)
        eng.set_tuning("budget", 1)
        eng.set_tuning("sort_cap_max", 4096)
        eng.set_tuning("sort_cap", 4096)
        with pytest.raises(RuntimeError):
            eng.evaluate(words, L.UNGATED)
        eng.set_tuning("sort_cap_max", 1 << 31)
        check(fx, rules, eng.evaluate(words, L.UNGATED), False, "dense/sort-after-limit")
    finally:
        eng.set_tuning("sort_cap_max", 1 << 31)
        reset(eng)


@pytest.mark.parametrize("dedup", [1, 3])
def test_giants_at_512(engines, dedup):
    """Seeds cut into sub-items after 512 wave-mode steps (the smallest value
    the autotuner tries) on the high-fan-out graph, where thousands of hub
    seeds become giants."""
    say("giants-512", dedup)
    eng, fx, g, rules, words = engine_for(engines, "hub_fanout_full")
    try:
        eng.set_tuning("dedup", dedup)
        eng.set_tuning("giant_after", 512)
        eng.set_tuning("budget", 64)
        check(fx, rules, eng.evaluate(words, L.UNGATED), False, "hub/giants%d" % dedup)
        assert eng.get_tuning("last_giants") > 0
        check(fx, rules, eng.evaluate(words, 0), True, "hub/giants%d-gated" % dedup)
    finally:
        reset(eng)


def same(a, b):
    return all(np.array_equal(a[f], b[f]) for f in FIELDS)


def test_serving_stream_alone_mixed_batch_golden():
    """The serving stream at its real shape: requests of 1-8 atoms over the
    resident high-fan-out graph, no work filter. The first 2,000 go through
    the server as a Poisson stream (mixed into batches that overlap on the
    device), 40 of them alone, and all 2,000 as one plain batch: every result
    is the same. The first 400 also match the reference programs."""
    say("serving stream")
    if "serving_stream_400" not in goldens.names():
        pytest.skip("fixture serving_stream_400 not generated (tests/golden/make_golden.py c5)")
    fx, g, rules400 = goldens.load("serving_stream_400")
    stream = goldens.generate("serving_stream", {"n_first": 2000})
    assert [W.rule_text(r) for r in stream.rules[:400]] == fx["rules_text"]
    words = L.rule_words_for_graph(stream.rules, g)
    eng = L.Engine.from_graph(g)
    try:
        t = time.time()
        ref = eng.evaluate(words, 0)   # reference semantics: body size only where support > 0
        say("  batch of %d: %.1fs" % (words.shape[0], time.time() - t))
        check(fx, rules400, ref[:400], True, "stream/batch")
        srv = S.Server(eng, max_batch=2048, max_in_flight=3, flags=0)
        try:
            alone = np.zeros(40, S.COMPLETION_DTYPE)
            for i in range(40):
                srv.submit(words[i:i + 1], np.array([i], np.uint64))
                c = srv.poll(20000)
                while c.shape[0] == 0:
                    c = srv.poll(20000)
                alone[i] = c[0]
            assert np.all(alone["batchSize"] == 1)
            rng = np.random.default_rng(5)
            arrivals = np.cumsum(rng.exponential(1.0 / 2000.0, words.shape[0]))
            t = time.time()
            r = S.run_stream(srv, words, arrivals)
            st = srv.stats()
        finally:
            srv.stop()
        mixed = r["completions"]
        say("  stream: %.1fs, %.0f rules/s, p50 %.1f ms, p99 %.1f ms, %d batches, mean batch %.1f" % (
            time.time() - t, r["rules_per_s"], np.percentile(r["latency_ms"], 50),
            np.percentile(r["latency_ms"], 99), st["batches"], mixed["batchSize"].mean()))
        assert st["completed"] == 40 + words.shape[0] and not st["device_failure"]
        assert same(alone["result"], ref[:40])
        assert same(mixed["result"], ref)
        assert (mixed["batchSize"] > 1).any()
        check(fx, rules400, mixed["result"][:400], True, "stream/served")
    finally:
        eng.close()

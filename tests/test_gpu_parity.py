"""GPU parity: libgpu_rules.so against the reference programs' golden
outputs and against the CPU oracle, through the C ABI.

Every count is an exact integer, so every comparison is exact.
"""
import json
import os
import sys
import tempfile

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import workloads as W
from oracle import Oracle

import goldens

pytestmark = pytest.mark.gpu

SMALL = [n for n in goldens.names() if n.startswith("tiny") or n in ("small_paths_2k", "dense_conf_300")]


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
        cache[name] = (L.Engine.from_graph(g), fx, g, rules)
    return cache[name]


def check(fx, rules, res, gated, tag):
    for i, r in enumerate(rules):
        exp = goldens.expected(fx, i, r)
        got = res[i]
        assert got["status"] == 0, (tag, i)
        assert got["support"] == exp["support"], (tag, i, "support", got["support"], exp, fx["rules_text"][i])
        assert got["headSize"] == exp["headSize"], (tag, i, "headSize")
        assert got["headSupport"] == exp["headSupport"], (tag, i, "headSupport")
        assert got["bodySize"] == exp["bodySize"], (tag, i, "bodySize", got["bodySize"], exp, fx["rules_text"][i])
        if not gated:
            assert got["bodyPairs"] == exp["bodyPairs"], (tag, i, "bodyPairs", got["bodyPairs"], exp,
                                                          fx["rules_text"][i])


@pytest.mark.parametrize("name", goldens.names())
def test_golden_batched(engines, name):
    say("batched", name)
    eng, fx, g, rules = engine_for(engines, name)
    words = L.rule_words_for_graph(rules, g)
    check(fx, rules, eng.evaluate(words, L.UNGATED), False, name + "/ungated")
    check(fx, rules, eng.evaluate(words, 0), True, name + "/gated")


@pytest.mark.parametrize("name", goldens.names())
def test_golden_per_rule(engines, name):
    say("per-rule", name)
    eng, fx, g, rules = engine_for(engines, name)
    words = L.rule_words_for_graph(rules, g)
    check(fx, rules, eng.evaluate(words, L.PER_RULE | L.UNGATED), False, name + "/per-rule")


@pytest.mark.parametrize("name", ["tiny_2", "small_paths_2k", "dense_conf_300", "dbpedia_shaped_5k"])
@pytest.mark.parametrize("dedup", [1, 3])
def test_forced_wave_mode(engines, name, dedup):
    """Budget 1: every seed with more than one candidate is escalated, so the
    wave-mode kernels (shared pool, work donation, ballot exit, head-object
    bitmaps in HBM rows (dedup 1) or in LDS (dedup 3)) carry the counts."""
    say("wave-mode", name, dedup)
    eng, fx, g, rules = engine_for(engines, name)
    words = L.rule_words_for_graph(rules, g)
    eng.set_tuning("budget", 1)
    eng.set_tuning("dedup", dedup)
    try:
        assert eng.get_tuning("pairs_dedup") == {1: 0, 3: 2}[dedup]
        res = eng.evaluate(words, L.UNGATED)
        assert eng.last_stats()["heavy_seeds"] > 0
        check(fx, rules, res, False, name + "/wave")
        eng.set_tuning("use_graph", 0)
        check(fx, rules, eng.evaluate(words, L.UNGATED), False, name + "/wave-eager")
        res = eng.evaluate(words, L.UNGATED | L.PER_RULE)
        check(fx, rules, res, False, name + "/wave-per-rule")
    finally:
        eng.set_tuning("budget", 4096)
        eng.set_tuning("dedup", 0)
        eng.set_tuning("use_graph", 1)


@pytest.mark.parametrize("seed", range(6))
def test_random_rules_vs_oracle(seed):
    """Random rules with constants, reflexive heads, repeated and
    disconnected atoms, 1-8 atoms, on small dense graphs with self-loops."""
    say("random", seed)
    g = W.tiny_random(9000 + seed, n_entities=10 + seed, n_predicates=3, n_triples=60 + 10 * seed)
    rng = np.random.default_rng(seed)
    rules = [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 8, n_vars=3 + k % 4, p_const=0.15,
                           closed=(k % 6 != 1), reflexive_head=(k % 13 == 3)) for k in range(120)]
    exp = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        for flags in (L.UNGATED, L.UNGATED | L.PER_RULE):
            for budget in (4096, 2):
                eng.set_tuning("budget", budget)
                res = eng.evaluate(words, flags)
                for i in range(len(rules)):
                    tag = (seed, flags, budget, i, W.rule_text(rules[i]))
                    assert res[i]["support"] == exp[i][0], tag
                    assert res[i]["headSize"] == exp[i][1], tag
                    assert res[i]["headSupport"] == exp[i][2], tag
                    assert res[i]["bodyPairs"] == exp[i][3], tag
                    assert res[i]["bodySize"] == exp[i][4], tag
    finally:
        eng.close()


@pytest.mark.parametrize("seed", range(4))
def test_single_frame_counts_vs_oracle(seed):
    """Rows of 16 entries and more on a dense graph with self-loops, so
    seeds whose last row is their only frame are counted without
    enumeration, and those rows hold the bound values the count must leave
    out (the seed itself, the intermediate variables). Rules of 1-3 atoms,
    some with constants, fused support on and off, every bitmap dedup mode,
    the batched and the per-rule path, against the oracle."""
    say("single-frame", seed)
    g = W.tiny_random(7100 + seed, n_entities=30 + seed, n_predicates=2, n_triples=1100 + 40 * seed)
    rng = np.random.default_rng(70 + seed)
    rules = [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 3, n_vars=2 + k % 3,
                           p_const=0.1 if k % 4 == 3 else 0.0, closed=True, reflexive_head=(k % 17 == 5))
             for k in range(90)]
    exp = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        for dedup, fuse, flags in ((0, 1, L.UNGATED), (1, 1, L.UNGATED), (3, 0, L.UNGATED), (0, 1, 0),
                                   (0, 1, L.UNGATED | L.PER_RULE)):
            eng.set_tuning("dedup", dedup)
            eng.set_tuning("fuse_support", fuse)
            res = eng.evaluate(words, flags)
            for i in range(len(rules)):
                tag = (seed, dedup, fuse, flags, i, W.rule_text(rules[i]))
                assert res[i]["support"] == exp[i][0], tag
                assert res[i]["bodySize"] == exp[i][4], tag
                if flags & L.UNGATED:
                    assert res[i]["bodyPairs"] == exp[i][3], tag
    finally:
        eng.close()


@pytest.mark.parametrize("name", ["tiny_0", "tiny_3", "small_paths_2k"])
def test_text_abi(name):
    """The reference entry points: TTL file in, rule lines and JSON objects
    in, 56-byte GpuRuleResult out."""
    say("text-abi", name)
    fx, g, rules = goldens.load(name)
    api = L.GpuRulesLib()
    with tempfile.TemporaryDirectory() as d:
        ttl = os.path.join(d, "g.ttl")
        W.write_ttl(g, ttl)
        ctx = api.load_graph(ttl)
        try:
            for i, r in enumerate(rules):
                exp = goldens.expected(fx, i, r)
                out = api.evaluate_rule(ctx, fx["rules_text"][i])
                assert (out.support, out.headSize, out.headSupport, out.bodySize) == (
                    exp["support"], exp["headSize"], exp["headSupport"], exp["bodySize"]), (name, i)
                hc = out.support / out.headSize if out.headSize else 0.0
                cf = out.support / out.bodySize if out.bodySize else 0.0
                assert out.headCoverage == hc and out.confidence == cf
                if i % 3 == 0:
                    outj = api.evaluate_rule_json(ctx, W.rule_json(r))
                    assert (outj.support, outj.bodySize) == (out.support, out.bodySize), (name, i)
            # malformed rule -> -1; unknown predicate -> rc 0 with zeros
            bad = L.GpuRuleResult()
            assert api.lib.gpu_evaluate_rule(ctx, b"( ?a <p> ?b", bad) == -1
            zero = api.evaluate_rule(ctx, "( ?a <https://ex.org/p/999> ?b ) => ( ?a <https://ex.org/p/0> ?b )")
            assert zero.support == 0 and zero.bodySize == 0
        finally:
            api.free_graph(ctx)


def test_graph_replay_is_stable():
    """A resident batch replays its captured HIP graph; every replay,
    interleaved with eager batches and host synchronisation, must agree with
    the eager result."""
    say("graph-replay")
    fx, g, rules = goldens.load("dbpedia_shaped_5k")
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        eng.set_tuning("use_graph", 0)
        eager = eng.evaluate(words, L.UNGATED).copy()
        check(fx, rules, eager, False, "eager")
        eng.set_tuning("use_graph", 1)
        b = eng.batch(words, L.UNGATED)
        for k in range(5):
            res = b.run()
            assert np.array_equal(res, eager), k
            other = eng.evaluate(words[: 50 + 10 * k], 0)       # eager-built batch in between
            check(fx, rules[: 50 + 10 * k], other, True, "interleaved")
        b.close()
    finally:
        eng.close()


@pytest.mark.skip(reason="cross-wave seed splitting is switched off: its launches ran past the time limit on the "
                         "GPU (tiny_5, dedup 1) twice without a cause shown; the wave kernels run without it")
@pytest.mark.parametrize("name", ["tiny_5", "small_paths_2k", "dense_conf_300", "dbpedia_shaped_5k"])
@pytest.mark.parametrize("dedup", [1, 2, 3])
def test_forced_cross_wave_split(engines, name, dedup):
    """split_after 1 and budget 1: wave-mode seeds are split across waves as
    soon as any wave is idle -- tasks, shared group rows (or sort keys),
    group found flags -- and every count stays exact."""
    say("split", name, dedup)
    eng, fx, g, rules = engine_for(engines, name)
    words = L.rule_words_for_graph(rules, g)
    eng.set_tuning("budget", 1)
    eng.set_tuning("dedup", dedup)
    eng.set_tuning("split_after", 1)
    try:
        splits = 0
        for flags, graph in ((L.UNGATED, 1), (L.UNGATED, 0), (0, 1), (L.UNGATED | L.PER_RULE, 1)):
            eng.set_tuning("use_graph", graph)
            res = eng.evaluate(words, flags)
            check(fx, rules, res, not (flags & L.UNGATED), "%s/split%d/%d" % (name, dedup, flags))
            splits += eng.get_tuning("last_splits")
        say("splits", splits)
        if name in ("dense_conf_300", "dbpedia_shaped_5k"):
            assert splits > 0
    finally:
        eng.set_tuning("budget", 4096)
        eng.set_tuning("dedup", 0)
        eng.set_tuning("split_after", 256)
        eng.set_tuning("use_graph", 1)


@pytest.mark.parametrize("name", ["tiny_5", "small_paths_2k", "dense_conf_300", "dbpedia_shaped_5k"])
@pytest.mark.parametrize("dedup", [1, 2, 3])
def test_forced_giant_seeds(engines, name, dedup):
    """giant_after 1 and budget 1: nearly every wave-mode seed is abandoned
    after one step and listed as a giant, cut into sub-items (level-0 or
    level-1 ranges) that all waves run against the giant's shared bitmap
    row, sort keys or found flag; every count stays exact. The per-rule
    launches cut giants too (per_rule_giant_work 0: for every rule)."""
    say("giants", name, dedup)
    eng, fx, g, rules = engine_for(engines, name)
    words = L.rule_words_for_graph(rules, g)
    eng.set_tuning("budget", 1)
    eng.set_tuning("dedup", dedup)
    eng.set_tuning("giant_after", 1)
    eng.set_tuning("per_rule_giant_work", 0)
    try:
        giants = 0
        for flags, graph in ((L.UNGATED, 1), (L.UNGATED, 0), (0, 1), (L.UNGATED, 1), (L.UNGATED | L.PER_RULE, 1),
                             (L.PER_RULE, 1)):
            eng.set_tuning("use_graph", graph)
            res = eng.evaluate(words, flags)
            check(fx, rules, res, not (flags & L.UNGATED), "%s/giant%d/%d" % (name, dedup, flags))
            giants += eng.get_tuning("last_giants")
        say("giants listed", giants)
        if name in ("dense_conf_300", "dbpedia_shaped_5k"):
            assert giants > 0
    finally:
        eng.set_tuning("budget", 4096)
        eng.set_tuning("dedup", 0)
        eng.set_tuning("giant_after", 2048)
        eng.set_tuning("use_graph", 1)
        eng.set_tuning("per_rule_giant_work", 10 ** 7)


@pytest.mark.parametrize("seed", range(3))
def test_giant_seeds_random_rules(seed):
    """Random rules with constants, reflexive heads, repeated and disconnected
    atoms: with every wave-mode seed turned into a giant, counts equal the
    oracle's."""
    say("giants-random", seed)
    g = W.tiny_random(9300 + seed, n_entities=12 + seed, n_predicates=3, n_triples=80 + 10 * seed)
    rng = np.random.default_rng(300 + seed)
    rules = [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 8, n_vars=3 + k % 4, p_const=0.15,
                           closed=(k % 6 != 1), reflexive_head=(k % 13 == 3)) for k in range(120)]
    exp = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        eng.set_tuning("budget", 1)
        eng.set_tuning("giant_after", 1)
        eng.set_tuning("per_rule_giant_work", 0)
        for dedup, flags in ((1, L.UNGATED), (3, L.UNGATED), (3, L.UNGATED | L.PER_RULE), (2, L.UNGATED | L.PER_RULE)):
            eng.set_tuning("dedup", dedup)
            res = eng.evaluate(words, flags)
            for i in range(len(rules)):
                tag = (seed, dedup, flags, i, W.rule_text(rules[i]))
                assert res[i]["support"] == exp[i][0], tag
                assert res[i]["bodyPairs"] == exp[i][3], tag
                assert res[i]["bodySize"] == exp[i][4], tag
    finally:
        eng.close()


def _variants(rng, rule, n_predicates, k):
    """k rules with rule's body up to renaming: another head predicate, the
    non-head variables renamed, the body atoms shuffled."""
    head, body = rule
    hv = {t[1] for t in (head[0], head[2]) if t[0] == "v"}
    out = []
    for j in range(k):
        ren = lambda t: t if t[0] == "c" or t[1] in hv else ("v", t[1] + "_%d" % j)
        b = [(ren(a[0]), a[1], ren(a[2])) for a in body]
        b = [b[q] for q in rng.permutation(len(b))]
        out.append(((head[0], int(rng.integers(n_predicates)), head[2]), b))
    return out


@pytest.mark.parametrize("seed", range(3))
def test_shared_bodies_vs_oracle(seed):
    """Rules with equal bodies up to renaming share one body-pair search
    (share_bodies): base rules plus variants with other head predicates,
    renamed variables and shuffled atoms -- some variants fused, some not.
    Every dedup mode, gated and ungated, sharing and fusion on and off:
    counts equal the oracle's."""
    say("shared-bodies", seed)
    g = W.tiny_random(9700 + seed, n_entities=16 + seed, n_predicates=3, n_triples=120 + 10 * seed)
    rng = np.random.default_rng(700 + seed)
    base = [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 4, n_vars=3 + k % 3, p_const=0.1)
            for k in range(50)]
    rules = []
    for r in base:
        rules.append(r)
        rules += _variants(rng, r, g.n_predicates, 2)
    exp = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        for dedup in (1, 2, 3):
            eng.set_tuning("dedup", dedup)
            for share in (1, 0):
                eng.set_tuning("share_bodies", share)
                for fuse in (1, 0):
                    eng.set_tuning("fuse_support", fuse)
                    for flags in (L.UNGATED, 0):
                        res = eng.evaluate(words, flags)
                        shared = eng.get_tuning("last_shared")
                        assert (shared > 0) if share else (shared == 0), (seed, dedup, share, fuse, flags, shared)
                        for i in range(len(rules)):
                            tag = (seed, dedup, share, fuse, flags, i, W.rule_text(rules[i]))
                            assert res[i]["support"] == exp[i][0], tag
                            assert res[i]["bodySize"] == exp[i][4], tag
                            if flags == L.UNGATED:
                                assert res[i]["bodyPairs"] == exp[i][3], tag
        # the traffic model counts a shared search once
        eng.set_tuning("share_bodies", 1)
        t_on = eng.traffic(words)
        eng.set_tuning("share_bodies", 0)
        t_off = eng.traffic(words)
        assert t_on[:, 1].sum() < t_off[:, 1].sum()
    finally:
        eng.set_tuning("share_bodies", 1)
        eng.set_tuning("fuse_support", 1)
        eng.close()


@pytest.mark.parametrize("name", ["dbpedia_shaped_5k", "dense_conf_300"])
def test_autotune_keeps_counts(name):
    """The autotuner times its candidate settings on a batch and keeps the
    fastest; every setting it keeps is one of its candidates, and counts
    under the chosen settings equal the goldens."""
    say("autotune", name)
    fx, g, rules = goldens.load(name)
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        log = eng.autotune(words, L.UNGATED)
        assert len(log) >= 20, len(log)
        assert eng.get_tuning("budget") in (64, 128, 256, 1024, 4096, 16384, 65536)
        assert eng.get_tuning("thread_waves_per_cu") in (8, 16, 24, 32)
        assert eng.get_tuning("chunks_per_wave") in (1, 2, 4, 8, 16)
        assert eng.get_tuning("wave_waves_per_cu") in (4, 8, 16)
        assert eng.get_tuning("giant_after") in (512, 1024, 2048, 4096, 8192)
        assert eng.get_tuning("use_graph") == 1
        check(fx, rules, eng.evaluate(words, L.UNGATED), False, name + "/autotuned")
        check(fx, rules, eng.evaluate(words, 0), True, name + "/autotuned-gated")
    finally:
        eng.close()


@pytest.mark.parametrize("seed", range(6))
def test_fused_support_vs_oracle(seed):
    """Support counted in the body-pair pass (fuse_support): rules over two
    predicates, so body atoms on the head predicate are common -- most cannot
    ground to the head triple (the rule is fused), some can (it keeps its
    support pass). Bitmap rows, sort-unique (never fused) and LDS bitmaps;
    thread mode, wave mode and giant sub-items; gated and ungated; fusion on
    and off: counts equal the oracle's."""
    say("fused-support", seed)
    g = W.tiny_random(9600 + seed, n_entities=14 + seed, n_predicates=2, n_triples=90 + 10 * seed)
    rng = np.random.default_rng(600 + seed)
    rules = [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % (5 if seed < 3 else 8), n_vars=3 + k % 4,
                           p_const=0.1) for k in range(150)]
    exp = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        for dedup in (1, 2, 3):
            eng.set_tuning("dedup", dedup)
            for budget, after in ((4096, 2048), (1, 2048), (1, 1)):
                eng.set_tuning("budget", budget)
                eng.set_tuning("giant_after", after)
                for fuse in (1, 0):
                    eng.set_tuning("fuse_support", fuse)
                    assert eng.get_tuning("fuse_active") == (1 if fuse and dedup != 2 else 0)
                    for flags in (L.UNGATED, 0):
                        res = eng.evaluate(words, flags)
                        fused = eng.get_tuning("last_fused")
                        if fuse and dedup != 2:
                            assert fused > 0, (seed, dedup, budget, after, flags)
                        else:
                            assert fused == 0, (seed, dedup, budget, after, flags)
                        for i in range(len(rules)):
                            tag = (seed, dedup, budget, after, fuse, flags, i, W.rule_text(rules[i]))
                            assert res[i]["support"] == exp[i][0], tag
                            assert res[i]["bodySize"] == exp[i][4], tag
                            if flags == L.UNGATED:
                                assert res[i]["bodyPairs"] == exp[i][3], tag
    finally:
        eng.set_tuning("fuse_support", 1)
        eng.close()


@pytest.mark.parametrize("name", ["tiny_4", "dense_conf_300", "dbpedia_shaped_5k"])
def test_snapshot_context_matches(name):
    """A context loaded from a snapshot of a term-id graph gives the golden
    counts, identical to the context it was saved from."""
    say("snapshot", name)
    fx, g, rules = goldens.load(name)
    words = L.rule_words_for_graph(rules, g)
    eng = L.Engine.from_graph(g)
    with tempfile.TemporaryDirectory() as d:
        snap = os.path.join(d, "g.snap")
        try:
            ref = eng.evaluate(words, L.UNGATED).copy()
            eng.save_snapshot(snap)
        finally:
            eng.close()
        eng2 = L.Engine.from_snapshot(snap)
        try:
            res = eng2.evaluate(words, L.UNGATED)
            check(fx, rules, res, False, name + "/snapshot")
            assert np.array_equal(res, ref)
        finally:
            eng2.close()


def test_snapshot_text_abi():
    """Turtle graph -> context -> snapshot -> context: the reference entry
    points resolve rule text to the same terms and give the same results."""
    say("snapshot-text-abi")
    fx, g, rules = goldens.load("small_paths_2k")
    api = L.GpuRulesLib()
    with tempfile.TemporaryDirectory() as d:
        ttl = os.path.join(d, "g.ttl")
        snap = os.path.join(d, "g.snap")
        W.write_ttl(g, ttl)
        ctx = api.load_graph(ttl)
        try:
            ref = [api.evaluate_rule(ctx, t) for t in fx["rules_text"]]
            assert api.lib.gpu_save_snapshot(ctx, snap.encode()) == 0
        finally:
            api.free_graph(ctx)
        ctx2 = api.lib.gpu_load_snapshot(snap.encode())
        assert ctx2
        try:
            for i, t in enumerate(fx["rules_text"]):
                out = api.evaluate_rule(ctx2, t)
                assert (out.support, out.headSize, out.headSupport, out.bodySize) == (
                    ref[i].support, ref[i].headSize, ref[i].headSupport, ref[i].bodySize), i
        finally:
            api.free_graph(ctx2)


def _snapshot_bytes(ctx_save, d, tag):
    p = os.path.join(d, tag + ".snap")
    ctx_save(p)
    return open(p, "rb").read()


@pytest.mark.parametrize("name", ["tiny_0", "tiny_5", "small_paths_2k", "dense_conf_300", "dbpedia_shaped_5k"])
def test_device_build_matches_host_build(name, monkeypatch):
    """The GPU-sorted build of a term-id graph (with every triple given
    twice, so the repeats must go) is byte-identical to the host build --
    compared through both graphs' snapshot files -- and gives the golden
    counts."""
    say("device-build", name)
    fx, g, rules = goldens.load(name)
    s = np.concatenate([g.s, g.s[::-1]]).astype(np.int32)
    p = np.concatenate([g.p, g.p[::-1]]).astype(np.int64) + g.n_entities
    o = np.concatenate([g.o, g.o[::-1]]).astype(np.int32)
    nt = g.n_entities + g.n_predicates
    words = L.rule_words_for_graph(rules, g)
    with tempfile.TemporaryDirectory() as d:
        snaps = {}
        for mode in ("0", "1"):
            monkeypatch.setenv("GPU_RULES_DEVICE_BUILD", mode)
            eng = L.Engine.from_triples(s, p, o, nt)
            try:
                snaps[mode] = _snapshot_bytes(eng.save_snapshot, d, "m" + mode)
                if mode == "1":
                    check(fx, rules, eng.evaluate(words, L.UNGATED), False, name + "/device-build")
            finally:
                eng.close()
        assert snaps["0"] == snaps["1"]


def test_device_build_turtle(monkeypatch):
    """gpu_load_graph with the device build: same graph (term ids, prefixes,
    arrays -- snapshot bytes) and same text-ABI results as the host build."""
    say("device-build-turtle")
    fx, g, rules = goldens.load("small_paths_2k")
    api = L.GpuRulesLib()
    with tempfile.TemporaryDirectory() as d:
        ttl = os.path.join(d, "g.ttl")
        W.write_ttl(g, ttl)
        snaps, outs = {}, {}
        for mode in ("0", "1"):
            monkeypatch.setenv("GPU_RULES_DEVICE_BUILD", mode)
            ctx = api.load_graph(ttl)
            try:
                p = os.path.join(d, "t" + mode + ".snap")
                assert api.lib.gpu_save_snapshot(ctx, p.encode()) == 0
                snaps[mode] = open(p, "rb").read()
                outs[mode] = [(r.support, r.headSize, r.headSupport, r.bodySize)
                              for r in (api.evaluate_rule(ctx, t) for t in fx["rules_text"][:40])]
            finally:
                api.free_graph(ctx)
        assert snaps["0"] == snaps["1"]
        assert outs["0"] == outs["1"]


@pytest.mark.parametrize("name", ["dbpedia_shaped_5k", "hub_fanout_20k", "dense_conf_300"])
def test_overlapped_phases_replay(name):
    """Graph replays of an ungated batch with the support and body-pair
    phases on two streams: every replay, interleaved with eager passes and
    replays of the sequential graph, equals the eager result. Support fusion
    is switched off on alternate pairs of replays (with it, these batches
    have few or no support chunks left to overlap), which also recaptures
    the graph."""
    say("overlap", name)
    fx, g, rules = goldens.load(name)
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        eng.set_tuning("use_graph", 0)
        eager = eng.evaluate(words, L.UNGATED).copy()
        check(fx, rules, eager, False, name + "/eager")
        eng.set_tuning("use_graph", 1)
        b = eng.batch(words, L.UNGATED)
        for k in range(8):
            eng.set_tuning("overlap", k % 2)
            eng.set_tuning("fuse_support", (k // 2) % 2)
            assert np.array_equal(b.run(), eager), (name, k)
            other = eng.evaluate(words[: 10 + 5 * k], 0)
            check(fx, rules[: 10 + 5 * k], other, True, name + "/interleaved")
        b.close()
    finally:
        eng.set_tuning("overlap", 1)
        eng.set_tuning("fuse_support", 1)
        eng.close()

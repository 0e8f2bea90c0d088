"""Both wave-mode engines (tuning key wave_engine: 0 the pool of entries,
1 the level-synchronous search) against the reference programs' golden
outputs and the CPU oracle, whichever engine is the library's default.

Budget 1 sends nearly every seed to wave mode; giant_after 1 turns nearly
every wave-mode seed into a giant cut into sub-items. Every count is an exact
integer: tolerance 0.
"""
import sys

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import workloads as W
from oracle import Oracle

import goldens
from test_gpu_parity import check

pytestmark = pytest.mark.gpu

ENGINES = [0, 1]


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


def defaults(eng):
    for k, v in (("budget", 4096), ("dedup", 0), ("giant_after", 2048), ("use_graph", 1),
                 ("per_rule_giant_work", 10 ** 7), ("wave_engine", 1)):
        eng.set_tuning(k, v)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("name", ["tiny_1", "tiny_5", "small_paths_2k", "dense_conf_300", "dbpedia_shaped_5k",
                                  "hub_fanout_20k"])
@pytest.mark.parametrize("dedup", [1, 2, 3])
def test_engine_wave_mode_goldens(engines, engine, name, dedup):
    """Budget 1 (nearly every seed in wave mode) under each dedup mode
    (HBM rows, sort-unique, LDS bitmaps): batched (graph replay and eager,
    gated and ungated) and per-rule."""
    say("engine", engine, name, dedup)
    eng, fx, g, rules = engine_for(engines, name)
    words = L.rule_words_for_graph(rules, g)
    eng.set_tuning("wave_engine", engine)
    eng.set_tuning("budget", 1)
    eng.set_tuning("dedup", dedup)
    try:
        assert eng.get_tuning("wave_engine") == engine
        for flags, graph in ((L.UNGATED, 1), (L.UNGATED, 0), (0, 1), (L.UNGATED | L.PER_RULE, 1)):
            eng.set_tuning("use_graph", graph)
            res = eng.evaluate(words, flags)
            check(fx, rules, res, not (flags & L.UNGATED), "%s/e%d/d%d/f%d/g%d" % (name, engine, dedup, flags, graph))
    finally:
        defaults(eng)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("name", ["tiny_5", "dense_conf_300", "dbpedia_shaped_5k", "hub_fanout_20k"])
@pytest.mark.parametrize("dedup", [1, 2, 3])
def test_engine_giants_goldens(engines, engine, name, dedup):
    """giant_after 1: wave-mode seeds become giants, cut into sub-items that
    every wave runs with the engine under test; per-rule launches cut them
    too (per_rule_giant_work 0)."""
    say("engine-giants", engine, name, dedup)
    eng, fx, g, rules = engine_for(engines, name)
    words = L.rule_words_for_graph(rules, g)
    eng.set_tuning("wave_engine", engine)
    eng.set_tuning("budget", 1)
    eng.set_tuning("dedup", dedup)
    eng.set_tuning("giant_after", 1)
    eng.set_tuning("per_rule_giant_work", 0)
    try:
        giants = 0
        for flags in (L.UNGATED, 0, L.UNGATED | L.PER_RULE):
            res = eng.evaluate(words, flags)
            check(fx, rules, res, not (flags & L.UNGATED), "%s/giants/e%d/d%d/f%d" % (name, engine, dedup, flags))
            giants += eng.get_tuning("last_giants")
        if name in ("dense_conf_300", "dbpedia_shaped_5k", "hub_fanout_20k"):
            assert giants > 0
    finally:
        defaults(eng)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("seed", range(4))
def test_engine_random_rules_vs_oracle(engine, seed):
    """Random rules with constants (conflict checks), reflexive heads,
    repeated and disconnected atoms (SCAN levels), 1-8 atoms (all three slot
    classes), in wave mode and as giants, fused support on and off."""
    say("engine-random", engine, seed)
    g = W.tiny_random(9500 + seed, n_entities=12 + 2 * seed, n_predicates=3, n_triples=90 + 15 * seed)
    rng = np.random.default_rng(500 + seed)
    rules = [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 8, n_vars=3 + k % 5, p_const=0.15,
                           closed=(k % 6 != 1), reflexive_head=(k % 13 == 3)) for k in range(128)]
    exp = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        eng.set_tuning("wave_engine", engine)
        eng.set_tuning("budget", 1)
        eng.set_tuning("per_rule_giant_work", 0)
        for dedup, giant_after, fuse, flags in ((1, 0, 1, L.UNGATED), (3, 0, 1, L.UNGATED), (2, 0, 1, L.UNGATED),
                                                (3, 1, 1, L.UNGATED), (1, 1, 0, L.UNGATED), (3, 0, 1, 0),
                                                (3, 1, 1, L.UNGATED | L.PER_RULE), (2, 1, 1, L.UNGATED | L.PER_RULE)):
            eng.set_tuning("dedup", dedup)
            eng.set_tuning("giant_after", giant_after)
            eng.set_tuning("fuse_support", fuse)
            res = eng.evaluate(words, flags)
            for i in range(len(rules)):
                tag = (engine, seed, dedup, giant_after, fuse, flags, i, W.rule_text(rules[i]))
                assert res[i]["status"] == 0, tag
                assert res[i]["support"] == exp[i][0], tag
                assert res[i]["bodySize"] == exp[i][4], tag
                if flags & L.UNGATED:
                    assert res[i]["bodyPairs"] == exp[i][3], tag
    finally:
        eng.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_per_rule_giants_then_plain_rules(engine):
    """Per-rule launches cut giant seeds only for rules of large estimated
    work. A rule evaluated right after one whose sub-items counted must not
    see their count: the per-rule count is the sum of two words, the second
    written by the sub-items, and both are cleared before every rule."""
    say("per-rule-mixed", engine)
    g = W.tiny_random(7101, n_entities=31, n_predicates=2, n_triples=1140)
    rng = np.random.default_rng(71)
    rules = [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 3, n_vars=2 + k % 3,
                           p_const=0.1 if k % 4 == 3 else 0.0, closed=True, reflexive_head=(k % 17 == 5))
             for k in range(90)]
    exp = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    hg = L.HostGraph(g)
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        work = hg.work(words)
        thr = float(np.median(work[work > 0]))
        cut = work >= thr
        assert cut.any() and (~cut & (work > 0)).any()
        eng.set_tuning("wave_engine", engine)
        eng.set_tuning("budget", 1)
        eng.set_tuning("giant_after", 1)
        eng.set_tuning("per_rule_giant_work", int(thr))
        res = eng.evaluate(words, L.UNGATED | L.PER_RULE)
        for i in range(len(rules)):
            tag = (engine, i, bool(cut[i]), W.rule_text(rules[i]))
            assert res[i]["support"] == exp[i][0], tag
            assert res[i]["bodyPairs"] == exp[i][3], tag
    finally:
        eng.close()
        hg.close()

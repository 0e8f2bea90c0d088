"""GPU parity of the sort-unique body-pair dedup path (wave-mode seeds emit
(seed, head object) keys to HBM; an LDS radix sort + unique count turns them
into distinct pairs) against the reference programs' golden outputs, the CPU
oracle and the bitmap-row path. Every count is an exact integer: tolerance 0.
"""
import sys

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import serving as S
from masterthesis_gpu_acceleration_amd import workloads as W
from oracle import Oracle

import goldens
from test_gpu_parity import check

pytestmark = pytest.mark.gpu

FIELDS = ("support", "headSize", "headSupport", "bodySize", "bodyPairs", "status")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b):
    return all(np.array_equal(a[f], b[f]) for f in FIELDS)


@pytest.mark.parametrize("name", ["tiny_1", "tiny_4", "small_paths_2k", "dense_conf_300", "dbpedia_shaped_5k"])
@pytest.mark.parametrize("budget", [4096, 1])
def test_sort_dedup_goldens(name, budget):
    """Forced sort-unique dedup, batched (graph replay and eager) and per
    rule; budget 1 sends nearly every seed through the wave-mode kernel."""
    say("sort-dedup", name, budget)
    fx, g, rules = goldens.load(name)
    eng = L.Engine.from_graph(g)
    try:
        eng.set_tuning("dedup", 2)
        eng.set_tuning("budget", budget)
        assert eng.get_tuning("sort_active") == 1
        words = L.rule_words_for_graph(rules, g)
        check(fx, rules, eng.evaluate(words, L.UNGATED), False, name + "/sort")
        if budget == 1:
            assert eng.dedup_stats()["keys_emitted"] > 0
        eng.set_tuning("use_graph", 0)
        check(fx, rules, eng.evaluate(words, L.UNGATED), False, name + "/sort-eager")
        check(fx, rules, eng.evaluate(words, 0), True, name + "/sort-gated")
        check(fx, rules, eng.evaluate(words, L.UNGATED | L.PER_RULE), False, name + "/sort-per-rule")
    finally:
        eng.close()


def test_sort_buffer_overflow_grows_and_reruns():
    """Key buffers started at 4096 keys: the dense stress rules emit far more,
    so the pass is redone with larger buffers, and the counts stay exact."""
    say("sort-overflow")
    fx, g, rules = goldens.load("dense_conf_300")
    eng = L.Engine.from_graph(g)
    try:
        eng.set_tuning("dedup", 2)
        eng.set_tuning("budget", 1)
        eng.set_tuning("sort_cap", 4096)
        assert eng.get_tuning("sort_cap") == 4096
        words = L.rule_words_for_graph(rules, g)
        res = eng.evaluate(words, L.UNGATED)
        st = eng.dedup_stats()
        say("emitted", st)
        assert st["keys_emitted"] > 4096 and st["sort_cap"] >= st["keys_emitted"]
        check(fx, rules, res, False, "overflow/batched")
        eng.set_tuning("sort_cap", 4096)
        check(fx, rules, eng.evaluate(words, L.UNGATED | L.PER_RULE), False, "overflow/per-rule")
    finally:
        eng.close()


@pytest.mark.parametrize("seed", range(3))
def test_sort_equals_rows_and_oracle(seed):
    """Random rules with constants, reflexive and disconnected atoms: the two
    dedup strategies agree with each other and with the oracle."""
    say("sort-vs-rows", seed)
    g = W.tiny_random(7100 + seed, n_entities=14 + seed, n_predicates=3, n_triples=90 + 10 * seed)
    rng = np.random.default_rng(100 + seed)
    rules = [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 8, n_vars=3 + k % 4, p_const=0.15,
                           closed=(k % 6 != 1), reflexive_head=(k % 13 == 3)) for k in range(120)]
    exp = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    eng = L.Engine.from_graph(g)
    try:
        words = L.rule_words_for_graph(rules, g)
        eng.set_tuning("budget", 2)
        out = {}
        for mode in (1, 2):
            eng.set_tuning("dedup", mode)
            out[mode] = eng.evaluate(words, L.UNGATED).copy()
        assert same(out[1], out[2])
        for i in range(len(rules)):
            tag = (seed, i, W.rule_text(rules[i]))
            assert out[2][i]["support"] == exp[i][0], tag
            assert out[2][i]["bodyPairs"] == exp[i][3], tag
            assert out[2][i]["bodySize"] == exp[i][4], tag
    finally:
        eng.close()


def test_sort_graph_replay_is_stable():
    """Replays of a batch captured with the sort-unique pass (histogram,
    scan, scatter, unique kernels inside the graph) match the eager result."""
    say("sort-replay")
    fx, g, rules = goldens.load("dbpedia_shaped_5k")
    eng = L.Engine.from_graph(g)
    try:
        eng.set_tuning("dedup", 2)
        eng.set_tuning("budget", 16)
        words = L.rule_words_for_graph(rules, g)
        eng.set_tuning("use_graph", 0)
        eager = eng.evaluate(words, L.UNGATED).copy()
        check(fx, rules, eager, False, "sort-eager")
        eng.set_tuning("use_graph", 1)
        b = eng.batch(words, L.UNGATED)
        for k in range(4):
            assert np.array_equal(b.run(), eager), k
            other = eng.evaluate(words[: 40 + 10 * k], 0)
            check(fx, rules[: 40 + 10 * k], other, True, "sort-interleaved")
        b.close()
    finally:
        eng.close()


def test_sort_serving_alone_and_mixed():
    """Serving slots with their own sort buffers: each result is the same
    alone and mixed into batches."""
    say("sort-serving")
    fx, g, rules = goldens.load("dbpedia_shaped_5k")
    eng = L.Engine.from_graph(g)
    try:
        eng.set_tuning("dedup", 2)
        eng.set_tuning("budget", 8)
        words = L.rule_words_for_graph(rules, g)
        ref = eng.evaluate(words, L.UNGATED).copy()
        check(fx, rules, ref, False, "sort-serving-ref")
        srv = S.Server(eng, max_batch=128, max_in_flight=3, flags=L.UNGATED)
        try:
            n = 20
            alone = np.zeros(n, S.COMPLETION_DTYPE)
            for i in range(n):
                srv.submit(words[i:i + 1], np.array([i], np.uint64))
                c = srv.poll(5000)
                while c.shape[0] == 0:
                    c = srv.poll(5000)
                alone[i] = c[0]
            srv.submit(words, np.arange(words.shape[0], dtype=np.uint64))
            mixed = np.zeros(words.shape[0], S.COMPLETION_DTYPE)
            got = 0
            while got < words.shape[0]:
                c = srv.poll(100)
                mixed[c["tag"].astype(np.int64)] = c
                got += c.shape[0]
            st = srv.stats()
        finally:
            srv.stop()
        assert not st["device_failure"]
        assert same(alone["result"], ref[:n])
        assert same(mixed["result"], ref)
    finally:
        eng.close()

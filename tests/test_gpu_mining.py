"""Support-pruned evaluation (gpu_evaluate_rules_min, gpu_batch_create_min)
and the level-wise miner on the GPU.

* Against the library's full evaluation of the same rules (itself checked
  against the reference programs' goldens here and in the parity tests):
  support, head figures and head coverage of every rule are unchanged; body
  pairs, body size and confidence are unchanged for the rules that meet the
  thresholds, -1 / 0 / 0 for the others -- batched (graph replay and eager),
  per rule, and as a resident batch replayed twice, under each dedup mode.
* Thresholds at a rule's exact support and head coverage pass it; one above
  does not.
* The miner driven by the GPU finds the same rules with the same counts as
  the miner driven by the CPU oracle.
Every count is an exact integer: tolerance 0.
"""
import sys

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import mining as M
from masterthesis_gpu_acceleration_amd import workloads as W

import goldens
from mining_eval import OracleEvaluator
from test_gpu_parity import check

pytestmark = pytest.mark.gpu

SAME = ("support", "headSize", "headSupport", "headCoverage", "status")


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


def expect_pruned(full, res, thr, tag):
    for f in SAME:
        assert np.array_equal(full[f], res[f]), (tag, f)
    ok = (full["status"] == 0) & (full["support"] >= thr)
    assert np.array_equal(res["bodyPairs"][ok], full["bodyPairs"][ok]), tag
    assert np.all(res["bodyPairs"][~ok & (full["status"] == 0)] == -1), tag
    assert np.array_equal(res["bodySize"], np.where(ok, full["bodySize"], 0)), tag
    assert np.array_equal(res["confidence"], np.where(ok, full["confidence"], 0.0)), tag


def thresholds_of(lib, head_size, ms, hc):
    return np.array([lib.gpu_host_support_threshold(int(h), int(ms), float(hc)) for h in head_size], np.int64)


@pytest.mark.parametrize("name", ["tiny_5", "small_paths_2k", "dbpedia_shaped_5k", "dense_conf_300",
                                  "hub_fanout_20k"])
@pytest.mark.parametrize("dedup", [1, 2, 3])
def test_evaluate_min_against_full_evaluation(engines, name, dedup):
    say("evaluate-min", name, dedup)
    eng, fx, g, rules = engine_for(engines, name)
    words = L.rule_words_for_graph(rules, g)
    eng.set_tuning("dedup", dedup)
    try:
        full = eng.evaluate(words, 0).copy()
        check(fx, rules, full, True, "%s/full/d%d" % (name, dedup))
        sup = full["support"][full["support"] > 0]
        med = int(np.median(sup)) if sup.size else 1
        for ms, hc in ((1, 0.0), (2, 0.0), (med, 0.0), (1, 0.05), (3, 0.2), (M.SUPPORT_ONLY, 0.0)):
            thr = thresholds_of(eng.lib, full["headSize"], ms, hc)
            for flags, graph in ((0, 1), (0, 0), (L.PER_RULE, 1)):
                eng.set_tuning("use_graph", graph)
                res = eng.evaluate_min(words, ms, hc, flags)
                expect_pruned(full, res, thr, (name, dedup, ms, hc, flags, graph))
            eng.set_tuning("use_graph", 1)
            b = eng.batch(words, 0, ms, hc)
            try:
                for rep in range(2):
                    expect_pruned(full, b.run(), thr, (name, dedup, ms, hc, "resident", rep))
            finally:
                b.close()
    finally:
        eng.set_tuning("dedup", 0)
        eng.set_tuning("use_graph", 1)


def test_threshold_boundaries(engines):
    say("boundaries")
    eng, fx, g, rules = engine_for(engines, "dbpedia_shaped_5k")
    words = L.rule_words_for_graph(rules, g)
    full = eng.evaluate(words, 0).copy()
    idx = [i for i in np.argsort(-full["support"], kind="stable")[:40] if full["support"][i] >= 2 and
           full["bodyPairs"][i] > 0][:12]
    assert len(idx) >= 6
    one = words[idx]
    for k, i in enumerate(idx):
        s, hs = int(full["support"][i]), int(full["headSize"][i])
        w = one[k:k + 1]
        assert eng.evaluate_min(w, s, 0.0)["bodyPairs"][0] == full["bodyPairs"][i]
        assert eng.evaluate_min(w, s + 1, 0.0)["bodyPairs"][0] == -1
        hc = s / hs
        assert eng.evaluate_min(w, 1, hc)["bodyPairs"][0] == full["bodyPairs"][i]
        assert eng.evaluate_min(w, 1, float(np.nextafter(hc, 2.0)))["bodyPairs"][0] == -1
    # the refused forms
    with pytest.raises(RuntimeError):
        eng.evaluate_min(one, 1, 0.0, L.UNGATED)
    with pytest.raises(RuntimeError):
        eng.evaluate_min(one, 1, 1.5)
    with pytest.raises(RuntimeError):
        eng.evaluate_min(one, 1, float("nan"))


def test_support_only_runs_no_pair_search(engines):
    """Rules whose support threshold exceeds every head triple cannot meet
    it: a support-only evaluation runs no body-pair search at all, and the
    supports are those of the full evaluation."""
    say("support-only")
    eng, fx, g, rules = engine_for(engines, "dbpedia_shaped_5k")
    words = L.rule_words_for_graph(rules, g)
    full = eng.evaluate(words, 0).copy()
    res = eng.evaluate_min(words, M.SUPPORT_ONLY, 0.0)
    assert eng.get_tuning("last_pair_searches") == 0
    assert np.array_equal(res["support"], full["support"])
    assert np.all(res["bodyPairs"][full["status"] == 0] == -1)
    res = eng.evaluate_min(words, 1, 0.0)
    assert eng.get_tuning("last_pair_searches") > 0
    expect_pruned(full, res, thresholds_of(eng.lib, full["headSize"], 1, 0.0), "support-only/after")


@pytest.mark.parametrize("case", [
    ("tiny", 2, 2, 3, 2, 0.05, 0.0),
    ("tiny", 3, 3, 3, 1, 0.0, 0.1),
    ("dbpedia_shaped_5k", 4, 8, 3, 2, 0.01, 0.0),
    ("small_paths_2k", 3, 6, 3, 3, 0.02, 0.0),
])
def test_miner_gpu_equals_oracle(case):
    src, nh, nb, max_atoms, ms, hc, mc = case
    say("miner", case)
    if src == "tiny":
        g = W.tiny_random(5100 + nh, n_entities=14, n_predicates=nb, n_triples=90)
    else:
        _, g, _ = goldens.load(src)
    heads, bodies = M.frequent_predicates(g, nh), M.frequent_predicates(g, nb)
    miner = M.Miner(heads, bodies, max_atoms, ms, hc, mc)
    eng = L.Engine.from_graph(g)
    try:
        got, st = miner.mine(M.GpuEvaluator(eng, g, miner))
    finally:
        eng.close()
    exp, st0 = miner.mine(OracleEvaluator(g, ms, hc))
    assert st.candidates == st0.candidates and st.passed == st0.passed
    assert M.results_table(got) == M.results_table(exp)
    assert len(got) > 0


def test_oversized_batches_split_in_halves(engines):
    """A one-shot evaluation whose phase seeds exceed the context's cap
    (forced low here with phase_seed_cap; the default is the heavy queue's
    32-bit range) runs in halves, recursively, with the same results; a
    resident batch over the cap is refused, and so is a cap below one
    rule's seeds."""
    say("split")
    eng, fx, g, rules = engine_for(engines, "dbpedia_shaped_5k")
    words = L.rule_words_for_graph(rules, g)
    info = np.zeros((len(rules), 8), np.int64)
    for i in range(len(rules)):
        assert eng.lib.gpu_host_plan_info(eng.ctx, words[i].ctypes.data, info[i].ctypes.data) == 0
    biggest = int(info[:, 5:7].max())
    total = int(info[:, 5:7].sum(axis=0).max())
    assert total > 4 * biggest
    full = [eng.evaluate(words, f).copy() for f in (L.UNGATED, 0)]
    pruned = eng.evaluate_min(words, 3, 0.02).copy()
    cap = eng.get_tuning("phase_seed_cap")
    try:
        eng.set_tuning("phase_seed_cap", biggest)
        for f, exp in zip((L.UNGATED, 0), full):
            res = eng.evaluate(words, f)
            check(fx, rules, res, not (f & L.UNGATED), "split/f%d" % f)
            for k in ("support", "bodyPairs", "bodySize", "headSize"):
                assert np.array_equal(res[k], exp[k]), k
        res = eng.evaluate_min(words, 3, 0.02)
        for k in ("support", "bodyPairs", "bodySize", "confidence"):
            assert np.array_equal(res[k], pruned[k]), k
        with pytest.raises(RuntimeError):
            eng.batch(words, 0)
        eng.set_tuning("phase_seed_cap", max(1, biggest // 2))
        with pytest.raises(RuntimeError):
            eng.evaluate(words, 0)
    finally:
        eng.set_tuning("phase_seed_cap", cap)
    check(fx, rules, eng.evaluate(words, 0), True, "split/after")

"""The level-wise miner (masterthesis_gpu_acceleration_amd.mining) on the CPU,
with the oracle as its evaluator: its output equals the exhaustive
enumeration of closed rules filtered by the thresholds, so the refinement
reaches every rule and the support / head-coverage pruning loses none.
Also the canonical rule keys and the library's support gate arithmetic
(gpu_host_support_threshold, no GPU call). Counts are exact: tolerance 0;
head coverage and confidence are the same IEEE quotients.
"""
import itertools
import math

import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import mining as M
from masterthesis_gpu_acceleration_amd import workloads as W

from mining_eval import OracleEvaluator, all_closed_rules, expected_table


def v(n):
    return ("v", n)


def test_canonical_ignores_atom_order_and_variable_names():
    head = (v("a"), 0, v("b"))
    r1 = (head, ((v("a"), 1, v("x")), (v("x"), 2, v("b"))))
    r2 = (head, ((v("y"), 2, v("b")), (v("a"), 1, v("y"))))
    r3 = (head, ((v("a"), 1, v("x")), (v("b"), 2, v("x"))))   # another orientation: another rule
    assert M.canonical(r1) == M.canonical(r2)
    assert M.canonical(r1) != M.canonical(r3)
    assert M.canonical(M.rule_from_key(M.canonical(r1))) == M.canonical(r1)
    # the head variables keep their roles: swapping them is another rule
    r4 = (head, ((v("b"), 1, v("x")), (v("x"), 2, v("a"))))
    assert M.canonical(r4) != M.canonical(r1)


def test_refinements_respect_the_closing_bound():
    head = (v("a"), 0, v("b"))
    # at two atoms of three, a rule with three variables occurring once cannot be closed
    r = (head, ((v("a"), 1, v("v2")),))
    for key in M.refinements(r, [0, 1], 2):
        assert M.is_closed(M.rule_from_key(key))          # the last atom must close the rule
    assert M.refinements((head, ((v("a"), 1, v("b")),)), [0], 1) == set()
    keys = M.refinements((head, ()), [0, 1], 3)
    assert len(keys) == 2 * 6 - 1    # six variable patterns per predicate, minus the head atom itself


@pytest.mark.parametrize("seed,n_pred,max_atoms,min_support,min_hc,min_conf", [
    (0, 2, 3, 1, 0.0, 0.0),
    (1, 2, 3, 2, 0.05, 0.0),
    (2, 2, 3, 4, 0.2, 0.1),
    (3, 3, 2, 1, 0.0, 0.0),
    (4, 3, 2, 3, 0.1, 0.0),
])
def test_miner_equals_exhaustive_enumeration(seed, n_pred, max_atoms, min_support, min_hc, min_conf):
    g = W.tiny_random(4400 + seed, n_entities=9 + seed, n_predicates=n_pred, n_triples=40 + 8 * seed)
    preds = list(range(n_pred))
    miner = M.Miner(preds, preds, max_atoms, min_support, min_hc, min_conf)
    ev = OracleEvaluator(g, min_support, min_hc)
    mined, st = miner.mine(ev)
    got = M.results_table(mined)
    exp, n_all = expected_table(g, preds, preds, max_atoms, min_support, min_hc, min_conf)
    assert got == exp
    assert len(st.candidates) == ev.calls <= max_atoms
    if min_support > 1 or min_hc > 0:   # pruning evaluates no more candidates than the unpruned loop
        _, st0 = M.Miner(preds, preds, max_atoms).mine(OracleEvaluator(g))
        assert sum(st.candidates) <= sum(st0.candidates)
    assert n_all >= len(exp)
    for k, (sup, hs, body, hc, conf) in got.items():
        assert sup >= min_support and hc >= min_hc and conf >= min_conf and body >= sup


def test_enumeration_counts_closed_rules():
    # one predicate, one atom: p(a,b) is the head itself, so only p(b,a)
    assert len(all_closed_rules([0], [0], 1)) == 1
    # two atoms over one predicate: the closed, connected bodies of two atoms
    keys = all_closed_rules([0], [0], 2)
    assert all(M.is_closed(M.rule_from_key(k)) for k in keys)
    assert M.canonical(((v("a"), 0, v("b")), ((v("b"), 0, v("a")),))) in keys
    assert M.canonical(((v("a"), 0, v("b")), ((v("a"), 0, v("c")), (v("c"), 0, v("b"))))) in keys


def test_support_threshold_arithmetic():
    f = L.load_library().gpu_host_support_threshold
    assert f(100, 1, 0.0) == 1 and f(100, 0, 0.0) == 1 and f(100, -5, 0.0) == 1
    assert f(100, 7, 0.0) == 7
    for hs in (1, 3, 7, 10, 97, 1000, 123457):
        for hc in (1e-9, 0.01, 0.1, 1 / 3, 0.3, 0.5, 0.999, 1.0):
            t = f(hs, 1, hc)
            # the smallest support whose double quotient reaches hc
            want = next((s for s in range(1, hs + 1) if s / hs >= hc), None)
            assert t == (want if want is not None else hs + 1), (hs, hc, t)
            assert f(hs, t + 3, hc) == t + 3
    assert f(0, 1, 0.1) == 2 ** 31 - 1
    assert f(10, 1, 1.5) == -1 and f(10, 1, -0.1) == -1 and f(10, 1, math.nan) == -1 and f(-1, 1, 0.0) == -1
    assert f(10, 2 ** 40, 0.0) == 2 ** 31 - 1


def test_words_of_keys_mean_the_same_rules():
    """The miner's rule words (variables by canonical number) and the
    structured rules' words (variables by first appearance) give the same
    counts."""
    from oracle import Oracle
    g = W.tiny_random(4471, n_entities=10, n_predicates=3, n_triples=60)
    keys = sorted(M.key_refinements((1, ()), [0, 1, 2], 3))
    keys += sorted(set().union(*(M.key_refinements(k, [0, 2], 3) for k in keys[:6])))
    o = Oracle(g.s, g.p, g.o)
    a = o.eval(M.words_of_keys(keys, 0))
    b = o.eval(W.encode_rules([M.rule_from_key(k) for k in keys]))
    assert (a[:, :5] == b[:, :5]).all()
    assert M.words_of_keys(keys, g.n_entities)[:, 2].tolist() == [k[0] + g.n_entities for k in keys]


@pytest.mark.parametrize("seed", range(3))
def test_native_refinement_equals_python(seed):
    """gpu_host_refine_rules (C++, host threads) against key_refinements, the
    same refinement written in Python, on random parents of 0-3 atoms; also
    closed_mask against key_closed and words_of_keys over records against
    the same over keys."""
    import random
    rnd = random.Random(seed)
    preds = [0, 2, 5, 7][: 2 + seed]
    level = sorted(M.key_refinements((3, ()), preds, 4))
    parents = [(3, ())] + rnd.sample(level, min(20, len(level)))
    lvl2 = sorted(set().union(*(M.key_refinements(k, preds, 4) for k in parents[1:6])))
    parents += rnd.sample(lvl2, min(40, len(lvl2)))
    for max_atoms in (2, 3, 4):
        want = sorted(set().union(*(M.key_refinements(k, preds, max_atoms) for k in parents)))
        got = M.refine(M.key_records(parents), preds, max_atoms)
        assert sorted(M.record_key(r) for r in got) == want
        assert got.shape[0] == len(want)   # sorted and without repeats
        assert M.closed_mask(got).tolist() == [M.key_closed(M.record_key(r)) for r in got]
        assert (M.words_of_keys(got, 11) == M.words_of_keys([M.record_key(r) for r in got], 11)).all()


def test_new_entry_points_refuse_bad_arguments():
    """The support-pruned and refinement entry points reject bad arguments
    before any device work (no GPU is touched: a NULL context fails first)."""
    import ctypes
    import numpy as np
    lib = L.load_library()
    out = np.zeros(4, L.RESULT64_DTYPE)
    w = np.zeros((1, 28), np.int32)
    assert lib.gpu_evaluate_rules_min(None, w.ctypes.data, 1, 0, 1, 0.0, out.ctypes.data) == -1
    assert not lib.gpu_batch_create_min(None, w.ctypes.data, 1, 0, 1, 0.0)
    rec = np.zeros((2, M.KEY_WORDS), np.int32)
    preds = np.array([0, 1], np.int32)
    buf = np.zeros((4, M.KEY_WORDS), np.int32)
    assert lib.gpu_host_refine_rules(rec.ctypes.data, -1, preds.ctypes.data, 2, 3, buf.ctypes.data, 4) == -1
    assert lib.gpu_host_refine_rules(rec.ctypes.data, 2, preds.ctypes.data, 2, 0, buf.ctypes.data, 4) == -1
    assert lib.gpu_host_refine_rules(rec.ctypes.data, 2, preds.ctypes.data, 2, 9, buf.ctypes.data, 4) == -1
    assert lib.gpu_host_refine_rules(rec.ctypes.data, 2, preds.ctypes.data, 2, 3, None, 4) == -1
    # room for none: the count alone (the caller then asks again with room)
    n = lib.gpu_host_refine_rules(rec.ctypes.data, 1, preds.ctypes.data, 2, 3, None, 0)
    assert n == len(M.key_refinements((0, ()), [0, 1], 3))
    assert lib.gpu_host_body_groundings(None, w.ctypes.data, 1, ctypes.c_void_p(0)) == -1

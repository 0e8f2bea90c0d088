"""gpu_host_body_groundings -- the pairs the reference's body-size kernel
emits before dedup (GPU_03/gpu_kernels.cuh:218-317: one pair per injective
body grounding) -- against brute force over every injective assignment of
the body's variables on tiny graphs, and the closed forms of one-atom and
two-atom path bodies. Exact integers: tolerance 0. No GPU call."""
import itertools

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import workloads as W


def brute(g, rule):
    head, body = rule
    names = sorted({t[1] for a in body for t in (a[0], a[2])})
    triples = set(zip(g.s.tolist(), g.p.tolist(), g.o.tolist()))
    n = 0
    for vals in itertools.permutations(range(g.n_entities), len(names)):
        m = dict(zip(names, vals))
        if all((m[a[0][1]], a[1], m[a[2][1]]) in triples for a in body):
            n += 1
    return n


def var_rules(rng, g, k):
    out = []
    while len(out) < k:
        r = W.random_rule(rng, g.n_entities, g.n_predicates, 1 + len(out) % 4, n_vars=2 + len(out) % 3, p_const=0.0,
                          closed=True)
        head, body = r
        vs = {t[1] for a in body for t in (a[0], a[2])}
        if head[0][1] in vs and head[2][1] in vs and head[0][1] != head[2][1]:
            out.append(r)
    return out


@pytest.mark.parametrize("seed", range(4))
def test_groundings_against_brute_force(seed):
    g = W.tiny_random(3300 + seed, n_entities=7, n_predicates=2 + seed % 2, n_triples=22 + 3 * seed)
    rules = var_rules(np.random.default_rng(seed), g, 40)
    hg = L.HostGraph(g)
    try:
        got = hg.body_groundings(L.rule_words_for_graph(rules, g))
    finally:
        hg.close()
    for i, r in enumerate(rules):
        assert got[i] == brute(g, r), (seed, i, W.rule_text(r))
    assert (got > 0).any()


def test_groundings_closed_forms_and_refusals():
    g = W.tiny_random(77, n_entities=30, n_predicates=2, n_triples=200)
    V = W.V
    s, p, o = g.s.astype(np.int64), g.p.astype(np.int64), g.o.astype(np.int64)
    # one atom p0(a, b): its triples that are not self-loops
    r1 = ((V("a"), 1, V("b")), [(V("a"), 0, V("b"))])
    n1 = int(((p == 0) & (s != o)).sum())
    # a path p0(a, c), p1(c, b): sum over c of out-degree x in-degree, minus the
    # pairs where a, b or c coincide
    r2 = ((V("a"), 0, V("b")), [(V("a"), 0, V("c")), (V("c"), 1, V("b"))])
    e0 = [(a, c) for a, c, q in zip(s, o, p) if q == 0 and a != c]
    e1 = {}
    for c, b, q in zip(s, o, p):
        if q == 1 and c != b:
            e1.setdefault(c, []).append(b)
    n2 = sum(1 for a, c in e0 for b in e1.get(c, []) if b != a)
    # constants, or a head variable missing from the body: no such pairs
    r3 = ((V("a"), 0, V("b")), [(V("a"), 1, W.C(3)), (V("a"), 0, V("b"))])
    r4 = ((V("a"), 0, V("b")), [(V("a"), 1, V("c"))])
    hg = L.HostGraph(g)
    try:
        got = hg.body_groundings(L.rule_words_for_graph([r1, r2, r3, r4], g))
    finally:
        hg.close()
    assert got.tolist() == [n1, n2, -1, -1]

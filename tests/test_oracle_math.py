"""Pins the CPU oracle by brute force and closed forms, independently of the
reference programs: exhaustive search over all variable assignments on tiny
graphs, and set/matrix identities for 1- and 2-atom rules."""
import itertools

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import workloads as W
from oracle import Oracle


def brute(g, rule):
    """Direct restatement of the definitions by exhaustive assignment:
    (support, set of distinct body pairs)."""
    T = set(zip(g.s.tolist(), g.p.tolist(), g.o.tolist()))
    head, body = rule
    vars_ = W.rule_vars(rule)
    ents = range(g.n_entities)

    def val(t, a):
        return a[t[1]] if t[0] == "v" else t[1]

    def body_ok(a, extra_used):
        trip = [(val(x[0], a), x[1], val(x[2], a)) for x in body]
        if any(t not in T for t in trip):
            return False
        used = trip + extra_used
        return len(set(used)) == len(used)

    hp = head[1]
    head_triples = sorted((s, o) for (s, p, o) in T if p == hp)
    support = 0
    bvars = [v for v in vars_ if v not in {t[1] for t in (head[0], head[2]) if t[0] == "v"}]
    for (s, o) in head_triples:
        a = {}
        ok = True
        for t, x in ((head[0], s), (head[2], o)):
            if t[0] == "c":
                ok &= t[1] == x
            elif t[1] in a:
                ok &= a[t[1]] == x
            else:
                a[t[1]] = x
        if not ok or len(set(a.values())) != len(a):
            continue
        found = False
        for combo in itertools.permutations([e for e in ents if e not in a.values()], len(bvars)):
            b = dict(a)
            b.update(zip(bvars, combo))
            if body_ok(b, [(s, hp, o)]):
                found = True
                break
        support += found
    pairs = set()
    hv = [t for t in (head[0], head[2])]
    inbody = {t[1] for x in body for t in (x[0], x[2]) if t[0] == "v"}
    if all(t[0] == "v" and t[1] in inbody for t in hv):
        bv = [v for v in vars_ if v in inbody]
        for combo in itertools.permutations(ents, len(bv)):
            b = dict(zip(bv, combo))
            if body_ok(b, []):
                pairs.add((b[hv[0][1]], b[hv[1][1]]))
    return support, pairs


@pytest.mark.parametrize("seed", range(8))
def test_oracle_equals_brute_force(seed):
    g = W.tiny_random(500 + seed, n_entities=5 + seed % 2, n_predicates=2, n_triples=18 + seed)
    rng = np.random.default_rng(seed)
    rules = []
    for k in range(40):
        rules.append(W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 4, n_vars=3 + k % 2,
                                   p_const=0.15, closed=(k % 5 != 0), reflexive_head=(k % 9 == 4)))
    res = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    for i, r in enumerate(rules):
        sup, pairs = brute(g, r)
        assert res[i][0] == sup, (i, r, res[i], sup)
        assert res[i][3] == len(pairs), (i, r, res[i], pairs)
        assert res[i][4] == (len(pairs) if sup > 0 else 0)


def _term_eq(a, b):
    """'never' / 'maybe' / 'always': can terms a and b take the same value
    under an injective binding (distinct variables never do; a variable may
    equal a constant)?"""
    if a[0] == "v" and b[0] == "v":
        return "always" if a[1] == b[1] else "never"
    if a[0] == "c" and b[0] == "c":
        return "always" if a[1] == b[1] else "never"
    return "maybe"


def _fusable(rule):
    """The library's condition for counting support in the body-pair pass:
    distinct head variables, both in the body, and no body atom on the head
    predicate that can ground to the head triple."""
    head, body = rule
    if head[0][0] != "v" or head[2][0] != "v" or head[0][1] == head[2][1]:
        return False
    inbody = {t[1] for x in body for t in (x[0], x[2]) if t[0] == "v"}
    if head[0][1] not in inbody or head[2][1] not in inbody:
        return False
    for x in body:
        if x[1] == head[1] and _term_eq(x[0], head[0]) != "never" and _term_eq(x[2], head[2]) != "never":
            return False
    return True


@pytest.mark.parametrize("seed", range(6))
def test_support_from_pairs_identity(seed):
    """For every fusable rule, support = #{distinct body pairs (s, o) with
    (s, p_head, o) a triple}, by exhaustive assignment; the identity fails
    for some rules the condition excludes (a body atom that can reuse the
    head triple), so the exclusion is needed."""
    g = W.tiny_random(700 + seed, n_entities=5 + seed % 2, n_predicates=2, n_triples=18 + seed)
    T = set(zip(g.s.tolist(), g.p.tolist(), g.o.tolist()))
    rng = np.random.default_rng(70 + seed)
    a, b, c = W.V("a"), W.V("b"), W.V("c")
    rules = [((a, 0, b), [(a, 0, b)]), ((a, 0, b), [(a, 0, c), (c, 1, b)]), ((a, 0, b), [(b, 0, a)])]
    rules += [W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 4, n_vars=3 + k % 2, p_const=0.15)
              for k in range(40)]
    fused = broken = 0
    for r in rules:
        sup, pairs = brute(g, r)
        hp = r[0][1]
        hits = sum((s, hp, o) in T for (s, o) in pairs)
        if _fusable(r):
            fused += 1
            assert sup == hits, (r, sup, hits)
        elif r[0][0][0] == "v" and r[0][2][0] == "v" and r[0][0][1] != r[0][2][1] and sup != hits:
            broken += 1
    assert fused >= 10
    assert broken >= 1   # (?a p ?b) <= (?a p ?b): support 0, every p pair a hit


@pytest.mark.parametrize("seed", range(4))
def test_pairs_do_not_depend_on_head_predicate(seed):
    """The distinct body pairs of a rule are the same set for any head
    predicate, any renaming of the non-head variables and any atom order
    (what lets rules with equal bodies share one search), by exhaustive
    assignment; swapping the head variables' roles transposes the set."""
    g = W.tiny_random(800 + seed, n_entities=5 + seed % 2, n_predicates=3, n_triples=20 + seed)
    rng = np.random.default_rng(80 + seed)
    checked = 0
    for k in range(30):
        r = W.random_rule(rng, g.n_entities, g.n_predicates, 1 + k % 3, n_vars=3 + k % 2, p_const=0.15)
        head, body = r
        if head[0][0] != "v" or head[2][0] != "v" or head[0][1] == head[2][1]:
            continue
        _, pairs = brute(g, r)
        hv = {head[0][1], head[2][1]}
        ren = lambda t: t if t[0] == "c" or t[1] in hv else ("v", t[1] + "_r")
        b2 = [(ren(a[0]), a[1], ren(a[2])) for a in body][::-1]
        for q in range(g.n_predicates):
            assert brute(g, ((head[0], q, head[2]), b2))[1] == pairs, (r, q)
        assert brute(g, ((head[2], head[1], head[0]), body))[1] == {(o, s) for (s, o) in pairs}
        checked += 1
    assert checked >= 10


def _rand_graph(seed, n=40, P=3, m=500):
    return W.tiny_random(seed, n_entities=n, n_predicates=P, n_triples=m)


@pytest.mark.parametrize("seed", range(4))
def test_closed_forms_one_atom(seed):
    g = _rand_graph(seed)
    T = {}
    for s, p, o in zip(g.s.tolist(), g.p.tolist(), g.o.tolist()):
        T.setdefault(p, set()).add((s, o))
    p, q = 0, 1
    a, b = W.V("a"), W.V("b")
    rules = [((a, p, b), [(a, q, b)]),          # same orientation, p != q
             ((a, p, b), [(b, q, a)]),          # inverse
             ((a, p, b), [(a, p, b)]),          # body atom == head atom
             ((a, p, b), [(a, q, b), (a, q, b)])]  # repeated body atom
    res = Oracle(g.s, g.p, g.o).eval(W.encode_rules(rules))
    pq = {(s, o) for (s, o) in T[p] & T[q] if s != o}
    inv = {(s, o) for (s, o) in T[p] if (o, s) in T[q] and s != o}
    q_pairs = {(s, o) for (s, o) in T[q] if s != o}
    assert res[0][0] == len(pq) and res[0][3] == len(q_pairs)
    assert res[1][0] == len(inv) and res[1][3] == len({(o, s) for (s, o) in q_pairs})
    assert res[2][0] == 0 and res[2][3] == len({(s, o) for (s, o) in T[p] if s != o})
    assert res[3][0] == 0 and res[3][3] == 0


@pytest.mark.parametrize("seed", range(4))
def test_closed_form_two_atom_path(seed):
    """bodyPairs of (?a q1 ?c) ^ (?c q2 ?b) = #{(a,b): a != b, exists c not in
    {a,b} with q1(a,c), q2(c,b)} -- a masked Boolean matrix product."""
    g = _rand_graph(100 + seed, n=30, P=2, m=400)
    n = g.n_entities
    A = np.zeros((2, n, n), np.int64)
    A[g.p, g.s, g.o] = 1
    Q1 = A[0].copy()
    Q2 = A[1].copy()
    np.fill_diagonal(Q1, 0)   # c != a
    np.fill_diagonal(Q2, 0)   # c != b
    M = Q1 @ Q2
    np.fill_diagonal(M, 0)    # a != b
    expect = int((M > 0).sum())
    a, b, c = W.V("a"), W.V("b"), W.V("c")
    rule = ((a, 0, b), [(a, 0, c), (c, 1, b)])
    res = Oracle(g.s, g.p, g.o).eval(W.encode_rules([rule]))
    assert res[0][3] == expect
    # support: head triples (a,b) in q1 with a != b and M[a,b] > 0, minus the
    # case where the witness path reuses the head triple: (a,0,c) == (a,0,b)
    # needs c == b, which the c-mask already excludes.
    H = A[0].copy()
    np.fill_diagonal(H, 0)
    assert res[0][0] == int(((H > 0) & (M > 0)).sum())

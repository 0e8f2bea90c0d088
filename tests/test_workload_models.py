"""CPU checks of the measurement and workload helpers: the path-work
estimates that bound the high-fan-out and serving generators agree with
brute-force path counting, the table form agrees with the sparse-product
form, and the host traffic model (the algorithmic-byte count behind the
roofline figure) gives the bytes its stated per-access costs give on a
hand-checked graph and is the same on every call (it runs on many threads)."""
import itertools

import numpy as np

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import workloads as W

import goldens


def brute_walks(g, steps):
    """Walks along `steps` by explicit enumeration (tiny graphs only)."""
    out = {}
    for p, s, o in zip(g.p, g.s, g.o):
        out.setdefault((int(p), 0, int(s)), []).append(int(o))
        out.setdefault((int(p), 1, int(o)), []).append(int(s))
    total = 0
    for start in range(g.n_entities):
        frontier = [start]
        for st in steps:
            frontier = [w for u in frontier for w in out.get((st >> 1, st & 1, u), [])]
        total += len(frontier)
    return total


def test_walk_matches_brute_force():
    g = W.tiny_random(3)
    pw = W._PathWork(g)
    rng = np.random.default_rng(0)
    for _ in range(40):
        k = int(rng.integers(1, 4))
        steps = [int(x) for x in rng.integers(0, 2 * g.n_predicates, k)]
        assert pw.walk(steps) == brute_walks(g, steps)


def test_work_tables_match_walks():
    g = W.small_paths(n_entities=3000, n_predicates=6, n_triples=12000, n_rules=2).graph
    pw = W._PathWork(g)
    rng = np.random.default_rng(1)
    heads = rng.integers(0, g.n_predicates, 60)
    steps = rng.integers(0, 2 * g.n_predicates, (60, 3))
    for L_ in (1, 2, 3):
        fast = pw.work_many(heads, steps, np.full(60, L_))
        for i in range(60):
            st = [int(x) for x in steps[i, :L_]]
            ref = W._rule_work(pw, int(heads[i]), st)
            assert abs(fast[i] - ref) <= 1e-4 * max(1.0, ref), (L_, i, fast[i], ref)
            assert abs(pw.work_fast(int(heads[i]), st) - ref) <= 1e-4 * max(1.0, ref)


def test_generators_respect_work_bound():
    g = W.small_paths(n_entities=3000, n_predicates=6, n_triples=12000, n_rules=2).graph
    pw = W._PathWork(g)
    t, reqs = W.serving_requests(graph=g, n_requests=400, rate_per_s=1000.0, max_work=2e5)
    assert len(reqs) == 400 and np.all(np.diff(t) > 0)
    lens = np.bincount([len(b) for _, b in reqs], minlength=9)
    assert lens[1:].min() > 20            # lengths stay spread over 1..8
    for head, body in reqs:
        L_ = min(len(body), 3)
        names = ["a"] + ["c%d" % j for j in range(1, L_)] + ["b"]
        assert W._rule_work(pw, head[1], W._path_steps(body[:L_], names)) <= 2e5 * 1.0001


def test_traffic_model_hand_case():
    # entities 0..3, predicates p=0, q=1 (terms 4, 5)
    s = np.array([0, 0, 1, 2], np.int32)
    p = np.array([0, 0, 0, 1], np.int32)
    o = np.array([1, 2, 3, 3], np.int32)
    g = W.Graph(s, p, o, 4, 2, "hand")
    h = L.HostGraph(g)
    # (?a p ?b) => (?a q ?b): support seeds are the q triples, one CHECK each
    rule = ((W.V("a"), 1, W.V("b")), [(W.V("a"), 0, W.V("b"))])
    tr = h.traffic(L.rule_words_for_graph([rule], g))
    # one q triple (2,3): seed 8 B; p-row of subject 2 is absent -> 16 B block read only
    assert tr[0, 0] == 8 + 16
    rule2 = ((W.V("a"), 0, W.V("b")), [(W.V("a"), 1, W.V("b"))])
    tr2 = h.traffic(L.rule_words_for_graph([rule2], g))
    # three p triples as seeds (8 B each); q-rows: subjects 0 and 1 have none (16 B each);
    # there is no p seed with subject 2, so no probe is made
    assert tr2[0, 0] == 3 * 8 + 3 * 16
    h.close()


def test_traffic_model_is_deterministic():
    fx, g, rules = goldens.load("dbpedia_shaped_5k")
    h = L.HostGraph(g)
    words = L.rule_words_for_graph(rules, g)
    a = h.traffic(words)
    b = h.traffic(words)
    h.close()
    assert np.array_equal(a, b)
    assert (a[:, 0] >= 8 * (a[:, 2] > 0)).all()
    assert a[:, 1].sum() > 0

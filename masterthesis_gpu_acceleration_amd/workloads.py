"""Seeded synthetic knowledge graphs and Horn rules for the counting path.

The reference's datasets (gpu-test-100k.ttl, dbpedia.ttl, original_train.ttl;
README.md:103-108) are not shipped with it, so every workload here is
generated from a fixed seed with numpy's PCG64 generator. A graph is three
int32 arrays (subject, predicate, object) over an entity index space
0..n_entities-1 and a predicate index space 0..n_predicates-1, already
deduplicated and free of self-loops unless a generator says otherwise.

Rules are plain Python structures: an atom is ``(subject, predicate, object)``
where subject/object are terms ``("v", name)`` or ``("c", entity_index)``.
``rule_text`` renders the reference's rule-line grammar
(GPU_03/RuleParser.cpp:251-322, e.g.
``( ?a <https://ex.org/p/3> ?c ) ^ ( ?c <https://ex.org/p/7> ?b ) => ( ?a <https://ex.org/p/1> ?b )``)
and ``write_ttl`` renders Turtle with full IRIs, one triple per line, which
the reference's Turtle reader (GPU_03/RdfIndexes.cpp:146-415) and this
project's reader both accept.

Configurations (sizes as named by the workload definitions):
  small_paths      20,000 entities, 20 predicates, 100,000 triples, 10 path rules
  dbpedia_shaped   50,000 entities, 300 predicates, 206,000 triples, 10,000 rules of 1-3 atoms
  hub_fanout       500,000 entities, 60 predicates, 6.8M triples, ~50 hubs, 50 rules of 2-4 atoms
  dense_conf       2,000 entities, 8 predicates, 100,000 triples, 32 rules of 3-8 atoms
  serving          Poisson request stream of rules (1-8 atoms) over the hub_fanout graph
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np

ENTITY_IRI = "https://ex.org/e/{}"
PRED_IRI = "https://ex.org/p/{}"

SEED_BASE = 0x5EED0000


@dataclass
class Graph:
    s: np.ndarray
    p: np.ndarray
    o: np.ndarray
    n_entities: int
    n_predicates: int
    name: str = ""

    @property
    def n_triples(self) -> int:
        return int(self.s.shape[0])


@dataclass
class Workload:
    graph: Graph
    rules: list = field(default_factory=list)
    name: str = ""


# ---------------------------------------------------------------- helpers

class _PathWork:
    """Estimates of the search work of path-shaped rules, used to keep
    generated workloads within a stated bound.

    Step (p, d) moves from a variable u to w over predicate p: d = 0 when u
    is the subject (w ranges over u's p-objects), d = 1 when u is the object.
    deg[2p+d][u] is the number of such moves from u. For a path of steps
    s1..sk from every start entity, the number of groundings (with repeats
    allowed, so an upper bound on the injective count) is
    1' A1 A2 ... Ak 1; work(x, s, y) = sum over moves u->w of step s of
    deg[x][u] * deg[y][w] gives the 3-step count in one pass over s's
    triples, and walk() gives exact path counts of any length by repeated
    sparse products."""

    def __init__(self, g: Graph):
        self.g = g
        n = g.n_entities
        P = g.n_predicates
        order = np.argsort(g.p, kind="stable")
        self.s = g.s[order].astype(np.int64)
        self.o = g.o[order].astype(np.int64)
        self.ptr = np.searchsorted(g.p[order], np.arange(P + 1))
        self.nnz = np.diff(self.ptr).astype(np.float64)
        self.deg = np.zeros((2 * P, n), np.float32)
        for p in range(P):
            a, b = self.ptr[p], self.ptr[p + 1]
            self.deg[2 * p] = np.bincount(self.s[a:b], minlength=n)
            self.deg[2 * p + 1] = np.bincount(self.o[a:b], minlength=n)
        self._dot = None
        self._work = {}

    def moves(self, step):
        p, d = step >> 1, step & 1
        a, b = self.ptr[p], self.ptr[p + 1]
        return (self.s[a:b], self.o[a:b]) if d == 0 else (self.o[a:b], self.s[a:b])

    def dot(self, x, y):
        if self._dot is None:
            self._dot = (self.deg.astype(np.float64) @ self.deg.T.astype(np.float64))
        return float(self._dot[x, y])

    def work(self, x, step, y):
        if step not in self._work:
            u, w = self.moves(step)
            self._work[step] = self.deg[:, u].astype(np.float64) @ self.deg[:, w].T.astype(np.float64)
        return float(self._work[step][x, y])

    def work_fast(self, head_pred, steps) -> float:
        """_rule_work for paths of at most 3 steps from the dot / work
        tables (no pass over the triples once the tables are built)."""
        k = len(steps)
        h = head_pred
        if k == 1:
            return float(max(self.nnz[steps[0] >> 1], self.nnz[h]))
        if k == 2:
            s1, s2 = steps
            return max(self.dot(s1 ^ 1, s2), self.dot(2 * h, s1), self.dot(2 * h + 1, s2 ^ 1))
        s1, s2, s3 = steps
        return max(self.work(s1 ^ 1, s2, s3), self.work(2 * h, s1, s2), self.work(2 * h + 1, s3 ^ 1, s2 ^ 1))

    def tables(self):
        """dot[x, y] and the full work[x, s, y] table (float32; one
        product per predicate, the reversed step is the transpose)."""
        if getattr(self, "_table", None) is None:
            P2 = self.deg.shape[0]
            tab = np.zeros((P2, P2, P2), np.float32)
            for p in range(P2 // 2):
                u, w = self.moves(2 * p)
                m = self.deg[:, u] @ self.deg[:, w].T
                tab[:, 2 * p, :] = m
                tab[:, 2 * p + 1, :] = m.T
            self._table = tab
            self._dot32 = (self.deg @ self.deg.T).astype(np.float64)
        return self._dot32, self._table

    def work_many(self, heads, steps, plen):
        """work_fast over arrays: heads[m], steps[m, 3] (only the first
        plen[i] used), plen[m] in 1..3."""
        dot, tab = self.tables()
        h = np.asarray(heads, np.int64)
        s1, s2, s3 = (steps[:, j].astype(np.int64) for j in range(3))
        w1 = np.maximum(self.nnz[s1 >> 1], self.nnz[h])
        w2 = np.maximum.reduce([dot[s1 ^ 1, s2], dot[2 * h, s1], dot[2 * h + 1, s2 ^ 1]])
        w3 = np.maximum.reduce([tab[s1 ^ 1, s2, s3], tab[2 * h, s1, s2], tab[2 * h + 1, s3 ^ 1, s2 ^ 1]])
        return np.where(plen == 1, w1, np.where(plen == 2, w2, w3.astype(np.float64)))

    def walk(self, steps, start=None):
        """Number of walks along `steps` (from every entity, or weighted by
        `start`), one sparse product per step."""
        v = np.ones(self.g.n_entities) if start is None else start.astype(np.float64)
        for st in steps:
            u, w = self.moves(st)
            v = np.bincount(w, weights=v[u], minlength=self.g.n_entities)
        return float(v.sum())


def _path_steps(body, names):
    """Steps of a path-shaped body walked from names[0]: atom (x p y) met at x
    is (p, 0), met at y is (p, 1). Atoms off the path (closing atoms) are
    left out."""
    steps = []
    left = list(body)
    cur = names[0]
    for nxt in names[1:]:
        for j, (x, p, y) in enumerate(left):
            if x == V(cur) and y == V(nxt):
                steps.append(2 * p)
                break
            if y == V(cur) and x == V(nxt):
                steps.append(2 * p + 1)
                break
        else:
            return None
        left.pop(j)
        cur = nxt
    return steps


def _rule_work(pw: _PathWork, head_pred: int, steps) -> float:
    """Largest of: body groundings from every start (the body-pair search),
    and the per-head-triple search from either end of the path (support)."""
    k = len(steps)
    h = head_pred
    if k == 1:
        return max(pw.nnz[steps[0] >> 1], pw.nnz[h])
    back = [s ^ 1 for s in reversed(steps)]
    pairs = pw.walk(steps)
    sup_a = pw.walk(steps[:-1], pw.deg[2 * h])
    sup_b = pw.walk(back[:-1], pw.deg[2 * h + 1])
    return max(pairs, sup_a, sup_b)


def _zipf_cdf(n: int, s: float) -> np.ndarray:
    w = np.arange(1, n + 1, dtype=np.float64) ** (-s)
    c = np.cumsum(w)
    return c / c[-1]


def _sample_cdf(rng, cdf: np.ndarray, size: int) -> np.ndarray:
    u = rng.random(size)
    return np.minimum(np.searchsorted(cdf, u, side="right"), cdf.shape[0] - 1)


def _dedup_keep_order(s, p, o, n_ent, n_pred, allow_loops=False):
    key = (s.astype(np.int64) * n_pred + p.astype(np.int64)) * n_ent + o.astype(np.int64)
    if not allow_loops:
        ok = s != o
        key = np.where(ok, key, -1)
    _, first = np.unique(key, return_index=True)
    first = np.sort(first)
    if not allow_loops:
        first = first[key[first] >= 0]
    return s[first], p[first], o[first]


def _fill_distinct(rng, target, n_ent, n_pred, draw, allow_loops=False):
    """Draw batches of (s,p,o) until `target` distinct triples remain; keep
    generation order so the result depends only on the seed."""
    S = np.empty(0, np.int32)
    P = np.empty(0, np.int32)
    O = np.empty(0, np.int32)
    while S.shape[0] < target:
        need = target - S.shape[0]
        s, p, o = draw(int(need * 1.3) + 64)
        S = np.concatenate([S, s.astype(np.int32)])
        P = np.concatenate([P, p.astype(np.int32)])
        O = np.concatenate([O, o.astype(np.int32)])
        S, P, O = _dedup_keep_order(S, P, O, n_ent, n_pred, allow_loops)
    return S[:target].copy(), P[:target].copy(), O[:target].copy()


class _Adjacency:
    """Incidence lists over a graph, used to sample satisfiable rule bodies."""

    def __init__(self, g: Graph):
        self.g = g
        n = g.n_entities
        self.out_order = np.argsort(g.s, kind="stable")
        self.out_ptr = np.zeros(n + 1, np.int64)
        np.add.at(self.out_ptr, g.s.astype(np.int64) + 1, 1)
        self.out_ptr = np.cumsum(self.out_ptr)
        self.in_order = np.argsort(g.o, kind="stable")
        self.in_ptr = np.zeros(n + 1, np.int64)
        np.add.at(self.in_ptr, g.o.astype(np.int64) + 1, 1)
        self.in_ptr = np.cumsum(self.in_ptr)
        self._pairs = None

    def incident(self, x):
        a = self.out_order[self.out_ptr[x]:self.out_ptr[x + 1]]
        b = self.in_order[self.in_ptr[x]:self.in_ptr[x + 1]]
        return a, b

    def pairs(self):
        if self._pairs is None:
            d = {}
            g = self.g
            for i in range(g.n_triples):
                d.setdefault((int(g.s[i]), int(g.o[i])), []).append(int(g.p[i]))
            self._pairs = d
        return self._pairs

    def step(self, rng, x, avoid):
        """Random triple incident to entity x whose other end is not in
        `avoid`. Returns (pred, other, x_is_subject) or None."""
        a, b = self.incident(x)
        na, nb = a.shape[0], b.shape[0]
        if na + nb == 0:
            return None
        for _ in range(8):
            k = int(rng.integers(na + nb))
            if k < na:
                t = a[k]
                other, xs = int(self.g.o[t]), True
            else:
                t = b[k - na]
                other, xs = int(self.g.s[t]), False
            if other not in avoid:
                return int(self.g.p[t]), other, xs
        return None


def V(name):
    return ("v", name)


def C(entity):
    return ("c", int(entity))


def _atom(x, pred, y, x_is_subject):
    return (x, pred, y) if x_is_subject else (y, pred, x)


# ------------------------------------------------------------ rendering

def term_text(t):
    if t[0] == "v":
        return "?" + t[1]
    return "<" + ENTITY_IRI.format(t[1]) + ">"


def atom_text(a):
    return "( {} <{}> {} )".format(term_text(a[0]), PRED_IRI.format(a[1]), term_text(a[2]))


def rule_text(rule) -> str:
    head, body = rule
    return " ^ ".join(atom_text(a) for a in body) + " => " + atom_text(head)


def _term_json(t):
    if t[0] == "v":
        return {"type": "variable", "value": "?" + t[1]}
    return {"type": "constant", "value": "<" + ENTITY_IRI.format(t[1]) + ">"}


def rule_json(rule) -> str:
    """RDFRules export shape (body_test_data/rules_to_json.py:9-12, 50-76)."""
    head, body = rule

    def aj(a):
        return {"subject": _term_json(a[0]), "predicate": "<" + PRED_IRI.format(a[1]) + ">",
                "object": _term_json(a[2])}

    return json.dumps({"body": [aj(a) for a in body], "head": aj(head)})


def write_ttl(g: Graph, path: str) -> None:
    with open(path, "w") as f:
        f.write("@prefix ex: <https://ex.org/> .\n")
        chunk = 1 << 16
        for i in range(0, g.n_triples, chunk):
            s = g.s[i:i + chunk]
            p = g.p[i:i + chunk]
            o = g.o[i:i + chunk]
            f.write("".join(
                "<https://ex.org/e/%d> <https://ex.org/p/%d> <https://ex.org/e/%d> .\n" % t
                for t in zip(s.tolist(), p.tolist(), o.tolist())))


def write_rules(rules, path: str) -> None:
    with open(path, "w") as f:
        for r in rules:
            f.write(rule_text(r) + "\n")


def write_rules_json(rules, path: str) -> None:
    with open(path, "w") as f:
        f.write("[\n" + ",\n".join(rule_json(r) for r in rules) + "\n]\n")


# ------------------------------------------------------------ small_paths

def small_paths(seed=SEED_BASE + 1, n_entities=20000, n_predicates=20, n_triples=100000,
                n_rules=10, zipf_s=1.1) -> Workload:
    """Predicate uniform, subject uniform, object Zipf(s) over a random entity
    ranking; path rules (?a p1 ?c1)...(?ck pk ?b) => (?a ph ?b), k in {2,3}."""
    rng = np.random.default_rng(seed)
    rank = rng.permutation(n_entities)
    cdf = _zipf_cdf(n_entities, zipf_s)

    def draw(k):
        return (rng.integers(0, n_entities, k), rng.integers(0, n_predicates, k),
                rank[_sample_cdf(rng, cdf, k)])

    s, p, o = _fill_distinct(rng, n_triples, n_entities, n_predicates, draw)
    g = Graph(s, p, o, n_entities, n_predicates, "small_paths")
    rules = []
    for _ in range(n_rules):
        k = int(rng.integers(2, 4))
        names = ["a"] + ["c%d" % i for i in range(1, k)] + ["b"]
        body = [(V(names[i]), int(rng.integers(n_predicates)), V(names[i + 1])) for i in range(k)]
        head = (V("a"), int(rng.integers(n_predicates)), V("b"))
        rules.append((head, body))
    return Workload(g, rules, "small_paths")


# ---------------------------------------------------------- dbpedia_shaped

_DB_SHAPES = {
    1: [[("a", "b")]],
    2: [[("a", "c"), ("c", "b")], [("a", "b"), ("a", "b")]],
    3: [[("a", "c"), ("c", "d"), ("d", "b")], [("a", "c"), ("c", "b"), ("a", "b")],
        [("a", "c"), ("c", "b"), ("c", "b")], [("a", "c"), ("a", "c"), ("c", "b")]],
}


def _graph_zipf_lognormal(rng, n_entities, n_predicates, n_triples, pred_s, sigma, obj_s):
    """Predicate frequency Zipf(pred_s); per predicate, subjects are drawn with
    a lognormal(0, sigma) out-degree; objects Zipf(obj_s) over a random
    ranking (uniform when obj_s == 0)."""
    pcdf = _zipf_cdf(n_predicates, pred_s)
    pcount = np.bincount(_sample_cdf(rng, pcdf, n_triples), minlength=n_predicates)
    rank = rng.permutation(n_entities)
    ocdf = _zipf_cdf(n_entities, obj_s) if obj_s > 0 else None
    S, P, O = [], [], []
    for p in range(n_predicates):
        need = int(pcount[p])
        if need == 0:
            continue
        deg = np.maximum(1, np.rint(rng.lognormal(0.0, sigma, need)).astype(np.int64))
        cum = np.cumsum(deg)
        nsub = int(np.searchsorted(cum, need) + 1)
        deg = deg[:nsub]
        deg[-1] -= int(cum[nsub - 1] - need)
        subj = rng.choice(n_entities, size=nsub, replace=nsub > n_entities)
        s = np.repeat(subj, deg)
        if ocdf is None:
            o = rng.integers(0, n_entities, need)
        else:
            o = rank[_sample_cdf(rng, ocdf, need)]
        S.append(s)
        P.append(np.full(need, p, np.int64))
        O.append(o)
    s = np.concatenate(S)
    p = np.concatenate(P)
    o = np.concatenate(O)
    perm = rng.permutation(s.shape[0])
    s, p, o = s[perm], p[perm], o[perm]
    s, p, o = _dedup_keep_order(s, p, o, n_entities, n_predicates)
    # top up to the exact triple count with uniform draws of the same shape
    if s.shape[0] < n_triples:
        def draw(k):
            pp = _sample_cdf(rng, pcdf, k)
            return rng.integers(0, n_entities, k), pp, (
                rng.integers(0, n_entities, k) if ocdf is None else rank[_sample_cdf(rng, ocdf, k)])
        extra = n_triples - s.shape[0]
        s2, p2, o2 = draw(extra * 2 + 64)
        s = np.concatenate([s, s2]); p = np.concatenate([p, p2]); o = np.concatenate([o, o2])
        s, p, o = _dedup_keep_order(s, p, o, n_entities, n_predicates)
    return s[:n_triples].astype(np.int32), p[:n_triples].astype(np.int32), o[:n_triples].astype(np.int32)


def _sample_guided_rule(rng, adj: _Adjacency, shape, pcdf, n_predicates, head_from_graph=0.6):
    """Ground a random walk of `shape` in the graph so the body is satisfiable,
    then pick the head predicate from a triple between the head endpoints when
    one exists (prob. head_from_graph), otherwise Zipf-random."""
    g = adj.g
    for _attempt in range(32):
        val = {}
        body = []
        t = int(rng.integers(g.n_triples))
        x, y = shape[0]
        xs = bool(rng.integers(2))
        if xs:
            val[x], val[y] = int(g.s[t]), int(g.o[t])
        else:
            val[x], val[y] = int(g.o[t]), int(g.s[t])
        if val[x] == val[y]:
            continue
        body.append(_atom(V(x), int(g.p[t]), V(y), xs))
        ok = True
        for (x, y) in shape[1:]:
            if x in val and y in val:
                preds = adj.pairs().get((val[x], val[y]), []) + \
                    [-1 - q for q in adj.pairs().get((val[y], val[x]), [])]
                if preds and rng.random() < 0.8:
                    q = preds[int(rng.integers(len(preds)))]
                    body.append((V(x), q, V(y)) if q >= 0 else (V(y), -1 - q, V(x)))
                else:
                    q = int(_sample_cdf(rng, pcdf, 1)[0])
                    body.append(_atom(V(x), q, V(y), bool(rng.integers(2))))
                continue
            src, dst = (x, y) if x in val else (y, x)
            st = adj.step(rng, val[src], set(val.values()))
            if st is None:
                ok = False
                break
            q, other, src_is_subject = st
            val[dst] = other
            body.append(_atom(V(src), q, V(dst), src_is_subject))
        if not ok:
            continue
        if rng.random() < head_from_graph:
            preds = adj.pairs().get((val["a"], val["b"]), [])
            hp = preds[int(rng.integers(len(preds)))] if preds else int(_sample_cdf(rng, pcdf, 1)[0])
        else:
            hp = int(_sample_cdf(rng, pcdf, 1)[0])
        return ((V("a"), hp, V("b")), body)
    # fall back to an unguided rule of the same shape
    body = [_atom(V(x), int(_sample_cdf(rng, pcdf, 1)[0]), V(y), bool(rng.integers(2))) for x, y in shape]
    return ((V("a"), int(_sample_cdf(rng, pcdf, 1)[0]), V("b")), body)


def _plant_templates(rng, n_entities, n_predicates, pcdf, n_templates, n_triples):
    """Rule regularities planted into a graph, as mined rules reflect them:
    each template is a closed body shape over {a,b,c,d} with fixed body and
    head predicates; each planted grounding adds the body triples and, with
    the template's confidence, the head triple. Returns (templates, s, p, o)."""
    templates = []
    for _ in range(n_templates):
        k = int(rng.choice([1, 2, 3], p=[0.15, 0.35, 0.5]))
        shapes = _DB_SHAPES[k]
        shape = shapes[int(rng.integers(len(shapes)))]
        preds = [int(x) for x in _sample_cdf(rng, pcdf, k)]
        orient = [bool(rng.integers(2)) for _ in range(k)]
        head = int(_sample_cdf(rng, pcdf, 1)[0])
        conf = float(rng.uniform(0.2, 0.9))
        templates.append((shape, preds, orient, head, conf))
    S, P, O = [], [], []
    per = max(1, n_triples // (n_templates * 3))
    for shape, preds, orient, head, conf in templates:
        for _ in range(per):
            val = {}
            for x in ("a", "b", "c", "d"):
                val[x] = int(rng.integers(n_entities))
            if len(set(val.values())) < 4:
                continue
            for (x, y), q, xs in zip(shape, preds, orient):
                a, b = (val[x], val[y]) if xs else (val[y], val[x])
                S.append(a); P.append(q); O.append(b)
            if rng.random() < conf:
                S.append(val["a"]); P.append(head); O.append(val["b"])
    return templates, np.asarray(S, np.int64), np.asarray(P, np.int64), np.asarray(O, np.int64)


def dbpedia_shaped(seed=SEED_BASE + 2, n_entities=50000, n_predicates=300, n_triples=206000,
                   n_rules=10000, atom_mix=(164, 1240, 8596), planted_share=0.3, n_templates=400,
                   template_rules=0.5) -> Workload:
    """Zipf(1.0) predicate frequencies, lognormal(0,1) out-degrees; closed
    rules of 1-3 body atoms over {?a,?b,?c,?d} in the DBpedia run's length mix
    (RunBenchmark_DBpedia.ipynb: 164 x 1, 1240 x 2, 8596 x 3 atoms).
    `planted_share` of the triples realise `n_templates` planted rule
    regularities (mined rules reflect such regularities); `template_rules`
    of the rules are drawn from the templates, the rest are grounded random
    walks with a head predicate taken from the graph where one connects the
    head variables."""
    rng = np.random.default_rng(seed)
    pcdf = _zipf_cdf(n_predicates, 1.0)
    n_planted = int(n_triples * planted_share)
    templates, ps, pp, po = _plant_templates(rng, n_entities, n_predicates, pcdf, n_templates, n_planted)
    s, p, o = _graph_zipf_lognormal(rng, n_entities, n_predicates, n_triples, 1.0, 1.0, 0.0)
    s = np.concatenate([ps, s.astype(np.int64)])
    p = np.concatenate([pp, p.astype(np.int64)])
    o = np.concatenate([po, o.astype(np.int64)])
    s, p, o = _dedup_keep_order(s, p, o, n_entities, n_predicates)
    s, p, o = s[:n_triples].astype(np.int32), p[:n_triples].astype(np.int32), o[:n_triples].astype(np.int32)
    g = Graph(s, p, o, n_entities, n_predicates, "dbpedia_shaped")
    adj = _Adjacency(g)
    tot = sum(atom_mix)
    counts = [int(round(n_rules * m / tot)) for m in atom_mix]
    counts[-1] = n_rules - sum(counts[:-1])
    by_len = {1: [], 2: [], 3: []}
    for t in templates:
        by_len[len(t[0])].append(t)
    rules = []
    for k, cnt in zip((1, 2, 3), counts):
        shapes = _DB_SHAPES[k]
        for _ in range(cnt):
            if by_len[k] and rng.random() < template_rules:
                shape, preds, orient, head, _conf = by_len[k][int(rng.integers(len(by_len[k])))]
                body = [_atom(V(x), q, V(y), xs) for (x, y), q, xs in zip(shape, preds, orient)]
                if rng.random() < 0.2:   # a variant: another head predicate
                    head = int(_sample_cdf(rng, pcdf, 1)[0])
                rules.append(((V("a"), head, V("b")), body))
            else:
                shape = shapes[int(rng.integers(len(shapes)))]
                rules.append(_sample_guided_rule(rng, adj, shape, pcdf, n_predicates))
    order = rng.permutation(len(rules))
    return Workload(g, [rules[i] for i in order], "dbpedia_shaped")


# ---------------------------------------------------------------- hub_fanout

def hub_fanout(seed=SEED_BASE + 3, n_entities=500000, n_predicates=60, n_triples=6800000,
               n_hubs=50, exponent=2.1, n_hub_preds=50, hub_obj_p=0.97, hub_subj_p=0.01, n_rules=50,
               min_atoms=2, max_atoms=4, max_path=3, max_work=1e9, tries=64) -> Workload:
    """Power-law degrees (Chung-Lu weights, exponent 2.1) for the tail, plus
    `n_hubs` hub entities: objects of the `n_hub_preds` hub predicates are a
    (uniformly chosen) hub with probability hub_obj_p, so every hub ends with
    a degree around 10^5 at the default sizes; rules of 2-4 atoms are paths
    grounded through a hub (at most `max_path` atoms long; longer rules add
    atoms that close over two path variables).

    A path through a hub with both neighbouring atoms on hub predicates has
    about deg(hub)^2 groundings per hub, and every further atom multiplies
    that by a degree-weighted fan-out, so most 3-4 atom walks have 10^10 to
    10^12 groundings. A rule is kept when its estimated search work
    (_rule_work: body groundings from every start, and the support search
    from either end) is at most `max_work`; up to `tries` walks are drawn per
    rule and the cheapest is kept when none qualifies."""
    rng = np.random.default_rng(seed)
    rank = rng.permutation(n_entities)
    hubs = rank[:n_hubs]
    alpha = 1.0 / (exponent - 1.0)
    wcdf = _zipf_cdf(n_entities - n_hubs, alpha)
    tail = rank[n_hubs:]
    hub_preds = np.arange(n_hub_preds)

    def draw(k):
        sp = tail[_sample_cdf(rng, wcdf, k)]
        pp = rng.integers(0, n_predicates, k)
        op = tail[_sample_cdf(rng, wcdf, k)]
        is_hub_pred = pp < n_hub_preds
        to_hub = is_hub_pred & (rng.random(k) < hub_obj_p)
        op = np.where(to_hub, hubs[rng.integers(0, n_hubs, k)], op)
        from_hub = is_hub_pred & ~to_hub & (rng.random(k) < hub_subj_p / max(1e-9, 1.0 - hub_obj_p))
        sp = np.where(from_hub, hubs[rng.integers(0, n_hubs, k)], sp)
        return sp, pp, op

    s, p, o = _fill_distinct(rng, n_triples, n_entities, n_predicates, draw)
    g = Graph(s, p, o, n_entities, n_predicates, "hub_fanout")
    adj = _Adjacency(g)
    pw = _PathWork(g)
    rules = []
    for i in range(n_rules):
        k = int(rng.integers(min_atoms, max_atoms + 1))
        plen = min(k, max_path)
        names = ["a"] + ["c%d" % j for j in range(1, plen)] + ["b"]
        shape = [(names[j], names[j + 1]) for j in range(plen)]
        best, best_w = None, None
        for _ in range(tries):
            r = _sample_hub_rule(rng, adj, shape, hub_preds, n_predicates, hubs, n_close=k - plen)
            steps = _path_steps(r[1], names)
            w = _rule_work(pw, r[0][1], steps) if steps is not None else float("inf")
            if best_w is None or w < best_w:
                best, best_w = r, w
            if w <= max_work:
                break
        rules.append(best)
    return Workload(g, rules, "hub_fanout")


def _sample_hub_rule(rng, adj, shape, hub_preds, n_predicates, hubs, n_close=0):
    """Path rule grounded through one hub: the walk starts at a hub and
    extends both ways, so at least one body atom sits on a hub predicate.
    `n_close` more atoms join two non-adjacent path variables, over a
    predicate that links their entities in the sampled grounding when one
    does."""
    g = adj.g
    hubset = set(int(h) for h in hubs)
    for _ in range(64):
        k = len(shape)
        mid = int(rng.integers(1, k)) if k > 1 else 0
        h = int(hubs[int(rng.integers(len(hubs)))])
        names = [shape[0][0]] + [y for _, y in shape]
        val = {names[mid]: h}
        atoms = [None] * k
        ok = True
        for j in range(mid - 1, -1, -1):        # walk left from the hub
            st = adj.step(rng, val[names[j + 1]], set(val.values()) | hubset)
            if st is None:
                ok = False
                break
            q, other, src_is_subject = st
            val[names[j]] = other
            atoms[j] = _atom(V(names[j + 1]), q, V(names[j]), src_is_subject)
        if not ok:
            continue
        for j in range(mid, k):                 # walk right
            st = adj.step(rng, val[names[j]], set(val.values()) | hubset)
            if st is None:
                ok = False
                break
            q, other, src_is_subject = st
            val[names[j + 1]] = other
            atoms[j] = _atom(V(names[j]), q, V(names[j + 1]), src_is_subject)
        if not ok:
            continue
        for _ in range(n_close):
            i, j = sorted(int(x) for x in rng.choice(len(names), size=2, replace=False))
            if j - i < 2 and len(names) > 2:
                i, j = 0, len(names) - 1
            atoms.append(_closing_atom(rng, adj, names[i], names[j], val, n_predicates))
        hp = int(hub_preds[int(rng.integers(len(hub_preds)))]) if rng.random() < 0.5 \
            else int(rng.integers(n_predicates))
        return ((V("a"), hp, V("b")), atoms)
    body = [(V(x), int(rng.integers(n_predicates)), V(y)) for x, y in shape]
    return ((V("a"), int(rng.integers(n_predicates)), V("b")), body)


def _closing_atom(rng, adj, x, y, val, n_predicates):
    """Atom between variables x and y: a predicate of a triple joining their
    sampled entities (either direction) if there is one, else a random one."""
    g = adj.g
    out_t, in_t = adj.incident(val[x])
    fwd = out_t[g.o[out_t] == val[y]]
    bwd = in_t[g.s[in_t] == val[y]]
    if fwd.shape[0] + bwd.shape[0]:
        k = int(rng.integers(fwd.shape[0] + bwd.shape[0]))
        if k < fwd.shape[0]:
            return (V(x), int(g.p[fwd[k]]), V(y))
        return (V(y), int(g.p[bwd[k - fwd.shape[0]]]), V(x))
    p = int(rng.integers(n_predicates))
    return (V(x), p, V(y)) if rng.random() < 0.5 else (V(y), p, V(x))


# ---------------------------------------------------------------- dense_conf

def dense_conf(seed=SEED_BASE + 4, n_entities=2000, n_predicates=8, n_triples=100000,
               n_rules=32, min_atoms=3, max_atoms=8) -> Workload:
    """Uniform subjects/objects (~50 average degree); path-like closed rules
    of 3-8 atoms with many groundings per rule."""
    rng = np.random.default_rng(seed)

    def draw(k):
        return (rng.integers(0, n_entities, k), rng.integers(0, n_predicates, k),
                rng.integers(0, n_entities, k))

    s, p, o = _fill_distinct(rng, n_triples, n_entities, n_predicates, draw)
    g = Graph(s, p, o, n_entities, n_predicates, "dense_conf")
    rules = []
    for i in range(n_rules):
        k = min_atoms + (i % (max_atoms - min_atoms + 1))
        rules.append(_dense_rule(rng, k, n_predicates))
    return Workload(g, rules, "dense_conf")


def _dense_rule(rng, k, n_predicates):
    """A path a -> ... -> b over k atoms when k <= 4; longer rules add closing
    atoms between already-bound path variables so the enumeration stays
    bounded while the depth grows."""
    plen = min(k, 3)
    names = ["a"] + ["c%d" % j for j in range(1, plen)] + ["b"]
    body = []
    for j in range(plen):
        body.append(_atom(V(names[j]), int(rng.integers(n_predicates)), V(names[j + 1]),
                          bool(rng.integers(2))))
    while len(body) < k:
        x, y = rng.choice(len(names), size=2, replace=False)
        body.append(_atom(V(names[x]), int(rng.integers(n_predicates)), V(names[y]),
                          bool(rng.integers(2))))
    return ((V("a"), int(rng.integers(n_predicates)), V("b")), body)


# ----------------------------------------------------------- tiny graphs

def tiny_random(seed, n_entities=8, n_predicates=3, n_triples=30, loops=True) -> Graph:
    """Small dense graph (self-loops allowed) for brute-force checks."""
    rng = np.random.default_rng(seed)

    def draw(k):
        return (rng.integers(0, n_entities, k), rng.integers(0, n_predicates, k),
                rng.integers(0, n_entities, k))

    total = n_entities * n_entities * n_predicates
    n_triples = min(n_triples, total if loops else total - n_entities * n_predicates)
    s, p, o = _fill_distinct(rng, n_triples, n_entities, n_predicates, draw, allow_loops=loops)
    return Graph(s, p, o, n_entities, n_predicates, "tiny")


def random_rule(rng, n_entities, n_predicates, n_atoms, n_vars=4, p_const=0.15,
                closed=True, reflexive_head=False):
    """Random rule over variables ?v0..?v{n_vars-1} with occasional constants,
    used to exercise the semantic corner cases (constants, reflexive atoms,
    repeated atoms, disconnected atoms)."""
    names = ["v%d" % i for i in range(n_vars)]

    def term():
        if rng.random() < p_const:
            return C(int(rng.integers(n_entities)))
        return V(names[int(rng.integers(n_vars))])

    body = [(term(), int(rng.integers(n_predicates)), term()) for _ in range(n_atoms)]
    if reflexive_head:
        head = (V("v0"), int(rng.integers(n_predicates)), V("v0"))
    else:
        hs = V("v0")
        ho = V("v1")
        r = rng.random()
        if r < 0.08:
            hs = C(int(rng.integers(n_entities)))
        elif r < 0.16:
            ho = C(int(rng.integers(n_entities)))
        head = (hs, int(rng.integers(n_predicates)), ho)
    if closed:
        used = {t[1] for a in body for t in (a[0], a[2]) if t[0] == "v"}
        for t in (head[0], head[2]):
            if t[0] == "v" and t[1] not in used and body:
                j = int(rng.integers(len(body)))
                a = body[j]
                body[j] = (t, a[1], a[2]) if rng.random() < 0.5 else (a[0], a[1], t)
                used = {t2[1] for a2 in body for t2 in (a2[0], a2[2]) if t2[0] == "v"}
    return (head, body)


def rule_vars(rule):
    head, body = rule
    out = []
    for a in body + [head]:
        for t in (a[0], a[2]):
            if t[0] == "v" and t[1] not in out:
                out.append(t[1])
    return out


def is_closed(rule) -> bool:
    """Every head variable appears in the body (AMIE closed-rule condition
    restricted to the head; the configurations only use such rules)."""
    head, body = rule
    bv = {t[1] for a in body for t in (a[0], a[2]) if t[0] == "v"}
    return all(t[0] != "v" or t[1] in bv for t in (head[0], head[2]))


# ----------------------------------------------------------- serving stream

def serving_requests(seed=SEED_BASE + 5, graph: Graph | None = None, n_requests=100000,
                     rate_per_s=20000.0, min_atoms=1, max_atoms=8, pred_s=1.0, max_work=None):
    """Poisson arrivals at `rate_per_s`; rule lengths uniform in
    [min_atoms, max_atoms]; predicates Zipf(pred_s). Returns (arrival_times_s,
    rules). A rule's body is a path a -> b of min(k, 3) atoms (random
    directions) plus atoms closing over two of its variables; the head is
    (?a p ?b).

    With `max_work` (needs `graph`), a path whose estimated search work
    (_PathWork.work_many) exceeds the bound is dropped and another path of
    the same length drawn, as a rule miner's pruning bound keeps such rules
    out of the stream; the lengths stay uniform. The kept fraction per path
    length is left in serving_requests.last_kept."""
    rng = np.random.default_rng(seed)
    n_pred = graph.n_predicates if graph is not None else 60
    pcdf = _zipf_cdf(n_pred, pred_s)
    gaps = rng.exponential(1.0 / rate_per_s, n_requests)
    t = np.cumsum(gaps)
    k = rng.integers(min_atoms, max_atoms + 1, n_requests)
    plen = np.minimum(k, 3)
    pw = _PathWork(graph) if (max_work is not None and graph is not None) else None
    pool = {}
    kept = {}
    for L in (1, 2, 3):
        need = int((plen == L).sum())
        got_p, got_f, got_h = [], [], []
        have, drawn = 0, 0
        while have < need:
            m = max(1024, 2 * (need - have))
            preds = _sample_cdf(rng, pcdf, 3 * m).reshape(m, 3)
            fwd = rng.integers(0, 2, (m, 3)).astype(bool)
            head = _sample_cdf(rng, pcdf, m)
            drawn += m
            if pw is not None:
                ok = pw.work_many(head, 2 * preds + (~fwd).astype(np.int64), np.full(m, L)) <= max_work
                preds, fwd, head = preds[ok], fwd[ok], head[ok]
            got_p.append(preds)
            got_f.append(fwd)
            got_h.append(head)
            have += preds.shape[0]
        if need:
            pool[L] = (np.concatenate(got_p), np.concatenate(got_f), np.concatenate(got_h))
            kept[L] = have / drawn
    used = {1: 0, 2: 0, 3: 0}
    rules = []
    for i in range(n_requests):
        L = int(plen[i])
        preds, fwd, head = pool[L]
        j = used[L]
        used[L] += 1
        names = ["a"] + ["c%d" % q for q in range(1, L)] + ["b"]
        body = [_atom(V(names[q]), int(preds[j, q]), V(names[q + 1]), bool(fwd[j, q])) for q in range(L)]
        while len(body) < int(k[i]):
            x, y = rng.choice(len(names), size=2, replace=False)
            body.append(_atom(V(names[x]), int(_sample_cdf(rng, pcdf, 1)[0]), V(names[y]),
                              bool(rng.integers(2))))
        rules.append(((V("a"), int(head[j]), V("b")), body))
    serving_requests.last_kept = kept
    return t, rules


def serving_stream(n_first=400, n_requests=100000, graph: Graph | None = None, seed=SEED_BASE + 5) -> Workload:
    """The first `n_first` requests of the serving stream -- `n_requests`
    Poisson arrivals, 1-8 atoms, predicates Zipf s=1.0, no work filter --
    over the high-fan-out graph (generated unless given), as a Workload. The
    requests a stream holds depend on its length (the paths of each length
    are drawn in pools sized for the whole stream), so the prefix is taken
    from the full-length stream; arrival rates only rescale the arrival times."""
    g = graph if graph is not None else hub_fanout().graph
    _, rules = serving_requests(seed=seed, graph=g, n_requests=n_requests, rate_per_s=1000.0, max_work=None)
    return Workload(g, rules[:n_first], "serving_stream")


# ------------------------------------------------------------ rule encoding

RULE_WORDS = 28   # [nBody, head(s,p,o), body[8](s,p,o)]


def encode_rules(rules, entity_id=None, pred_id=None) -> np.ndarray:
    """Structured rules -> int32 records of RULE_WORDS words, the layout of
    the library's GpuRuleDesc and of the oracle's input. A term >= 0 is a
    constant term id; a variable is -(k+1) with k its first-appearance
    number (body atoms subject-then-object, then the head, as
    GPU_03/RuleParser.cpp:251-322 numbers them). `entity_id`/`pred_id` map
    generator indices to term ids (identity when None)."""
    out = np.zeros((len(rules), RULE_WORDS), np.int32)
    for i, (head, body) in enumerate(rules):
        vmap = {}

        def enc(t):
            if t[0] == "v":
                if t[1] not in vmap:
                    vmap[t[1]] = len(vmap)
                return -(vmap[t[1]] + 1)
            return int(t[1]) if entity_id is None else int(entity_id[t[1]])

        def encp(q):
            return int(q) if pred_id is None else int(pred_id[q])

        out[i, 0] = len(body)
        for j, a in enumerate(body[:8]):
            s = enc(a[0])
            o = enc(a[2])
            out[i, 4 + 3 * j:7 + 3 * j] = (s, encp(a[1]), o)
        hs = enc(head[0])
        ho = enc(head[2])
        out[i, 1:4] = (hs, encp(head[1]), ho)
    return out

"""Level-wise rule mining over the library's support-pruned batches.

The reference's library serves RDFRules, an AMIE-style miner, through a JNA
binding (GPU_03/GpuRulesLib.scala:46-58). The miner refines rules one body
atom at a time and drops a rule once its support falls below minSupport or
its head coverage (support / headSize) below minHeadCoverage; support never
grows when an atom is added, so no refinement of a dropped rule can pass
(the reference restates the same threshold test for minSupport = 1 at
Rewritten_CPP/SupportCounting.cpp:67-90). `Miner` runs that loop over the
device: all candidates of one body length form one batch of
`gpu_evaluate_rules_min`, which counts every candidate's support exactly and
reports body pairs (for confidence) only for the candidates that meet the
thresholds, searching them only for those (a rule whose support the
body-pair pass counts itself gets both from one search).

Rules are the structured rules of `workloads` -- (head, body), atoms (s, p, o)
with terms ("v", name) -- over variables only: the head is p(?a, ?b), and a
refinement adds either a dangling atom (one variable of the rule and a fresh
one) or a closing atom (two variables of the rule), in either orientation,
over any body predicate. Atoms are never reflexive and never repeat the head
atom or a body atom. A rule is reported when it is closed (every variable
occurs at least twice in head and body) and meets the thresholds and
minConfidence. Candidates that no sequence of the remaining atoms could close
are not generated (each atom can close at most two variables that occur
once). Rules with constants are not generated.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field

import numpy as np

HEAD_S, HEAD_O = "a", "b"


def _fresh(k):
    return "v%d" % k


def var_counts(rule):
    head, body = rule
    cnt = {}
    for a in (head,) + tuple(body):
        for t in (a[0], a[2]):
            cnt[t[1]] = cnt.get(t[1], 0) + 1
    return cnt


def is_closed(rule) -> bool:
    return all(c >= 2 for c in var_counts(rule).values())


def canonical(rule):
    """A key equal for rules that differ only in the order of their body
    atoms and the names of their non-head variables: over the numberings of
    the non-head variables (2, 3, ...), the smallest sorted tuple of body
    atoms (predicate, subject number, object number)."""
    head, body = rule
    others = sorted({t[1] for a in body for t in (a[0], a[2])} - {HEAD_S, HEAD_O})
    best = None
    for perm in itertools.permutations(range(2, 2 + len(others))):
        names = dict(zip(others, perm))
        names[HEAD_S] = 0
        names[HEAD_O] = 1
        enc = tuple(sorted((p, names[s[1]], names[o[1]]) for s, p, o in body))
        if best is None or enc < best:
            best = enc
    return (head[1], best)


def rule_from_key(key):
    """The rule a canonical key stands for (variables a, b, v2, v3, ...)."""
    hp, enc = key

    def name(k):
        return HEAD_S if k == 0 else (HEAD_O if k == 1 else _fresh(k))

    body = tuple((("v", name(s)), p, ("v", name(o))) for p, s, o in enc)
    return ((("v", HEAD_S), hp, ("v", HEAD_O)), body)


def refinements(rule, body_predicates, max_atoms):
    """Rules with one more body atom (dangling or closing), as canonical keys;
    none whose variables that occur once could not all be closed by the
    atoms still allowed."""
    head, body = rule
    if len(body) >= max_atoms:
        return set()
    cnt = var_counts(rule)
    vars_ = sorted(cnt)
    fresh = _fresh(len(vars_))
    present = set(body) | {head}
    left = max_atoms - len(body) - 1   # atoms still allowed after this one
    out = set()
    pairs = [(x, y) for x in vars_ for y in vars_ if x != y]            # closing
    pairs += [(x, fresh) for x in vars_] + [(fresh, x) for x in vars_]  # dangling
    for x, y in pairs:
        c2 = dict(cnt)
        c2[x] = c2.get(x, 0) + 1
        c2[y] = c2.get(y, 0) + 1
        if sum(1 for c in c2.values() if c == 1) > 2 * left:
            continue
        for p in body_predicates:
            a = (("v", x), int(p), ("v", y))
            if a in present:
                continue
            out.add(canonical((head, tuple(body) + (a,))))
    return out


@dataclass
class MinedRule:
    rule: tuple
    support: int
    head_size: int
    body_size: int
    head_coverage: float
    confidence: float


@dataclass
class MiningStats:
    candidates: list = field(default_factory=list)   # per body length
    passed: list = field(default_factory=list)       # candidates meeting the thresholds, per length
    eval_s: list = field(default_factory=list)       # evaluator time per length
    total_s: float = 0.0


class Miner:
    """Level-wise refinement with support / head-coverage pruning.

    evaluate(rules) -> structured array with fields support, headSize,
    bodyPairs (-1 where not counted), headCoverage, confidence for the
    thresholds given to it; `GpuEvaluator` is the device's."""

    def __init__(self, head_predicates, body_predicates, max_atoms=3, min_support=1, min_head_coverage=0.0,
                 min_confidence=0.0):
        self.heads = [int(p) for p in head_predicates]
        self.bodies = [int(p) for p in body_predicates]
        self.max_atoms = int(max_atoms)
        self.min_support = max(1, int(min_support))
        self.min_hc = float(min_head_coverage)
        self.min_conf = float(min_confidence)

    def passes(self, r) -> bool:
        return r["support"] >= self.min_support and r["headCoverage"] >= self.min_hc and r["support"] > 0

    def mine(self, evaluate):
        st = MiningStats()
        t0 = time.perf_counter()
        level = set()
        for h in self.heads:
            head = (("v", HEAD_S), h, ("v", HEAD_O))
            level |= refinements((head, ()), self.bodies, self.max_atoms)
        out = []
        while level:
            keys = sorted(level)
            rules = [rule_from_key(k) for k in keys]
            t1 = time.perf_counter()
            res = evaluate(rules)
            st.eval_s.append(time.perf_counter() - t1)
            st.candidates.append(len(rules))
            nxt = set()
            npass = 0
            for rule, r in zip(rules, res):
                if int(r["status"]) != 0 or not self.passes(r):
                    continue
                npass += 1
                if is_closed(rule) and r["bodyPairs"] >= 0 and r["confidence"] >= self.min_conf:
                    out.append(MinedRule(rule, int(r["support"]), int(r["headSize"]), int(r["bodySize"]),
                                         float(r["headCoverage"]), float(r["confidence"])))
                nxt |= refinements(rule, self.bodies, self.max_atoms)
            st.passed.append(npass)
            level = nxt
        st.total_s = time.perf_counter() - t0
        out.sort(key=lambda m: canonical(m.rule))
        return out, st


# A minimum support no rule reaches: the batch counts support only.
SUPPORT_ONLY = 1 << 62


class GpuEvaluator:
    """Evaluates a level's candidates as support-pruned device batches: the
    closed candidates with the miner's thresholds (body pairs for those that
    meet them), the open ones for support only (their confidence is never
    reported, and their body pairs -- head variables joined by no atom --
    can be a cross product)."""

    def __init__(self, engine, graph, miner: Miner, flags: int = 0):
        from . import lib as L
        self.L = L
        self.engine = engine
        self.graph = graph
        self.miner = miner
        self.flags = flags
        self.device_ms = []   # per call: device time of its batches (gpu_last_batch_stats)

    def __call__(self, rules):
        words = self.L.rule_words_for_graph(rules, self.graph)
        closed = np.array([is_closed(r) for r in rules], bool)
        out = np.zeros(len(rules), self.L.RESULT64_DTYPE)
        ms = 0.0
        if closed.any():
            out[closed] = self.engine.evaluate_min(words[closed], self.miner.min_support, self.miner.min_hc,
                                                   self.flags)
            ms += self.engine.last_stats()["total_ms"]
        if (~closed).any():
            out[~closed] = self.engine.evaluate_min(words[~closed], SUPPORT_ONLY, 0.0, self.flags)
            ms += self.engine.last_stats()["total_ms"]
        self.device_ms.append(ms)
        return out


def results_table(mined):
    """Canonical key -> (support, headSize, bodySize, headCoverage, confidence)."""
    return {canonical(m.rule): (m.support, m.head_size, m.body_size, m.head_coverage, m.confidence)
            for m in mined}


def frequent_predicates(graph, k):
    """The k predicates with the most triples (ties by index)."""
    cnt = np.bincount(np.asarray(graph.p, np.int64), minlength=graph.n_predicates)
    order = sorted(range(graph.n_predicates), key=lambda q: (-int(cnt[q]), q))
    return [q for q in order[:k] if cnt[q] > 0]

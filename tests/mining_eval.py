"""Helpers of the mining tests: the CPU oracle as a mining evaluator, and an
exhaustive enumeration of the rules a level-wise miner must find.

OracleEvaluator returns what gpu_evaluate_rules_min returns, from the
oracle's full counts: support, head figures and head coverage for every
rule, body pairs only for the rules that meet the thresholds (-1 for the
others). Head coverage is the same IEEE double quotient the library reports.
"""
import itertools

import numpy as np

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import mining as M
from masterthesis_gpu_acceleration_amd import workloads as W
from oracle import Oracle


class OracleEvaluator:
    def __init__(self, g, min_support=1, min_head_coverage=0.0):
        self.oracle = Oracle(g.s, g.p, g.o)
        self.min_support = max(1, int(min_support))
        self.min_hc = float(min_head_coverage)
        self.calls = 0

    def __call__(self, keys):
        """Support of every rule (key records as the miner passes them, or
        canonical keys); body pairs of the closed rules that meet the
        thresholds (the miner reads no other rule's)."""
        self.calls += 1
        if isinstance(keys, np.ndarray):
            keys = [M.record_key(r) for r in keys]
        rules = [M.rule_from_key(k) for k in keys]
        words = W.encode_rules(rules)
        exp = self.oracle.eval_support(words)
        sup, hs = exp[:, 0], exp[:, 1]
        hc0 = np.where(hs > 0, sup.astype(np.float64) / np.maximum(hs, 1).astype(np.float64), 0.0)
        full = (sup >= self.min_support) & (hc0 >= self.min_hc) & np.array([M.key_closed(k) for k in keys], bool)
        if full.any():
            exp[full] = self.oracle.eval(words[full])
        sup, hs, hsup, pairs = exp[:, 0], exp[:, 1], exp[:, 2], exp[:, 3]
        pairs = np.where(full, pairs, -1)
        out = np.zeros(len(rules), L.RESULT64_DTYPE)
        out["support"] = sup
        out["headSize"] = hs
        out["headSupport"] = hsup
        hc = np.where(hs > 0, sup.astype(np.float64) / np.maximum(hs, 1).astype(np.float64), 0.0)
        out["headCoverage"] = hc
        ok = (sup >= self.min_support) & (hc >= self.min_hc)
        out["bodyPairs"] = np.where(ok, pairs, -1)
        body = np.where(ok & (sup > 0) & (pairs > 0), pairs, 0)
        out["bodySize"] = body
        out["confidence"] = np.where(body > 0, sup.astype(np.float64) / np.maximum(body, 1).astype(np.float64), 0.0)
        return out


def connected(rule) -> bool:
    """Every variable reachable from the head's through shared atoms."""
    head, body = rule
    parent = {}

    def find(x):
        while parent.setdefault(x, x) != x:
            x = parent[x]
        return x

    for a in (head,) + tuple(body):
        parent[find(a[0][1])] = find(a[2][1])
    root = find(head[0][1])
    return all(find(a[0][1]) == root and find(a[2][1]) == root for a in body)


def all_closed_rules(head_predicates, body_predicates, max_atoms):
    """Canonical keys of every closed, connected rule p(?a, ?b) <= body with
    1..max_atoms body atoms over variables, no reflexive atom, no repeated
    atom and no copy of the head atom: every set of distinct atoms over a
    pool of max_atoms + 2 variables, filtered (the miner builds its rules
    atom by atom instead)."""
    names = [M.HEAD_S, M.HEAD_O] + ["w%d" % k for k in range(max_atoms)]
    keys = set()
    for h in head_predicates:
        head = (("v", M.HEAD_S), int(h), ("v", M.HEAD_O))
        atoms = [(("v", x), int(p), ("v", y)) for x in names for y in names if x != y for p in body_predicates]
        for k in range(1, max_atoms + 1):
            for body in itertools.combinations(atoms, k):
                if head in body:
                    continue
                rule = (head, body)
                if M.is_closed(rule) and connected(rule):
                    keys.add(M.canonical(rule))
    return keys


def expected_table(g, head_predicates, body_predicates, max_atoms, min_support, min_hc, min_conf):
    """The miner's expected output by exhaustive enumeration and full counts."""
    keys = sorted(all_closed_rules(head_predicates, body_predicates, max_atoms))
    res = OracleEvaluator(g, min_support, min_hc)(keys)
    table = {}
    for k, r in zip(keys, res):
        sup = int(r["support"])
        if sup < max(1, min_support) or r["headCoverage"] < min_hc or r["confidence"] < min_conf:
            continue
        table[k] = (sup, int(r["headSize"]), int(r["bodySize"]), float(r["headCoverage"]), float(r["confidence"]))
    return table, len(keys)

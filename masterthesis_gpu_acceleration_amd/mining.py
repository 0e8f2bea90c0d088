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

The loop runs on key records (the canonical key of a rule as int32 words);
each level's refinements come from the library's host threads
(gpu_host_refine_rules), and `key_refinements` is the same step in Python,
which the tests hold the native one to.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field

import numpy as np

HEAD_S, HEAD_O = "a", "b"

# Inside the loop a rule is its canonical key (head predicate, atoms): the
# body atoms as (predicate, subject, object) with variables numbered 0 (the
# head subject), 1 (the head object), 2, 3, ... -- integers only, so the
# candidates of a level never become nested rule tuples.


def _fresh(k):
    return "v%d" % k


def _key_atoms(rule):
    """(head predicate, body atoms over variable numbers) of a structured rule."""
    head, body = rule
    names = {head[0][1]: 0, head[2][1]: 1}
    atoms = []
    for s, p, o in body:
        for t in (s, o):
            if t[1] not in names:
                names[t[1]] = len(names)
        atoms.append((int(p), names[s[1]], names[o[1]]))
    return head[1], atoms


def canonical_atoms(hp, atoms):
    """The canonical key of a body given as (predicate, subject, object)
    atoms over variable numbers: over the numberings of the non-head
    variables (2, 3, ...), the smallest sorted tuple of atoms."""
    others = sorted({v for a in atoms for v in (a[1], a[2]) if v > 1})
    if not others:
        return (hp, tuple(sorted(atoms)))
    best = None
    for perm in itertools.permutations(range(2, 2 + len(others))):
        m = dict(zip(others, perm))
        enc = tuple(sorted((p, m.get(s, s), m.get(o, o)) for p, s, o in atoms))
        if best is None or enc < best:
            best = enc
    return (hp, best)


def canonical(rule):
    """A key equal for rules that differ only in the order of their body
    atoms and the names of their non-head variables."""
    hp, atoms = _key_atoms(rule)
    return canonical_atoms(hp, atoms)


def rule_from_key(key):
    """The rule a canonical key stands for (variables a, b, v2, v3, ...)."""
    hp, enc = key

    def name(k):
        return HEAD_S if k == 0 else (HEAD_O if k == 1 else _fresh(k))

    body = tuple((("v", name(s)), p, ("v", name(o))) for p, s, o in enc)
    return ((("v", HEAD_S), hp, ("v", HEAD_O)), body)


def _counts(enc):
    cnt = {0: 1, 1: 1}
    for _, s, o in enc:
        cnt[s] = cnt.get(s, 0) + 1
        cnt[o] = cnt.get(o, 0) + 1
    return cnt


def key_closed(key) -> bool:
    return all(c >= 2 for c in _counts(key[1]).values())


def is_closed(rule) -> bool:
    """Every variable occurs at least twice in head and body."""
    return key_closed(canonical(rule))


def key_refinements(key, body_predicates, max_atoms):
    """Canonical keys of the rules with one more body atom (dangling or
    closing); none whose variables that occur once could not all be closed
    by the atoms still allowed."""
    hp, enc = key
    if len(enc) >= max_atoms:
        return set()
    cnt = _counts(enc)
    vars_ = sorted(cnt)
    fresh = len(vars_)
    present = set(enc)
    left = max_atoms - len(enc) - 1   # atoms still allowed after this one
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
            a = (int(p), x, y)
            if a in present or (x == 0 and y == 1 and a[0] == hp):   # a repeated atom, or the head atom
                continue
            out.add(canonical_atoms(hp, enc + (a,)))
    return out


def refinements(rule, body_predicates, max_atoms):
    """key_refinements of a structured rule."""
    return key_refinements(canonical(rule), body_predicates, max_atoms)


# A key record: the library's layout of a canonical key (GPU_RULE_KEY_WORDS
# int32: head predicate, atom count, then (predicate, subject, object) per
# atom, the rest zero), the form gpu_host_refine_rules reads and writes.
KEY_WORDS = 26
_MAX_ATOMS = 8


def key_records(keys):
    """Key records of canonical keys (hp, atoms)."""
    out = np.zeros((len(keys), KEY_WORDS), np.int32)
    for i, (hp, enc) in enumerate(keys):
        out[i, 0] = hp
        out[i, 1] = len(enc)
        for j, a in enumerate(enc):
            out[i, 2 + 3 * j:5 + 3 * j] = a
    return out


def record_key(rec):
    """The canonical key (hp, atoms) of one key record."""
    n = int(rec[1])
    return (int(rec[0]), tuple((int(rec[2 + 3 * j]), int(rec[3 + 3 * j]), int(rec[4 + 3 * j])) for j in range(n)))


def _atoms_used(K):
    """The largest atom count of the key records (0 for none)."""
    return int(min(_MAX_ATOMS, max(0, int(K[:, 1].max())))) if K.shape[0] else 0


def words_of_keys(keys, pred_offset=0):
    """GpuRuleDesc words (RULE_WORDS int32 each) of key records (or of
    canonical keys); predicate p becomes term p + pred_offset, variable k
    becomes -(k + 1)."""
    K = keys if isinstance(keys, np.ndarray) else key_records(keys)
    n = K.shape[0]
    na = _atoms_used(K)   # atom slots past every record's count stay zero
    A = K[:, 2:2 + 3 * na].reshape(n, na, 3)
    m = np.arange(na)[None, :] < K[:, 1:2]
    W = np.zeros((n, 28), np.int32)
    W[:, 0] = K[:, 1]
    W[:, 1] = -1
    W[:, 2] = K[:, 0] + pred_offset
    W[:, 3] = -2
    B = W[:, 4:4 + 3 * na].reshape(n, na, 3)
    B[:, :, 0] = np.where(m, -(A[:, :, 1] + 1), 0)
    B[:, :, 1] = np.where(m, A[:, :, 0] + pred_offset, 0)
    B[:, :, 2] = np.where(m, -(A[:, :, 2] + 1), 0)
    return W


def closed_mask(K):
    """Per key record: every variable occurs at least twice in head and body
    (each row's variable occurrences sorted: none may differ from both
    neighbours)."""
    n = K.shape[0]
    na = _atoms_used(K)
    A = K[:, 2:2 + 3 * na].reshape(n, na, 3)
    m = np.arange(na)[None, :] < K[:, 1:2]
    occ = np.concatenate([np.zeros((n, 1), np.int32), np.ones((n, 1), np.int32),
                          np.where(m, A[:, :, 1], -1), np.where(m, A[:, :, 2], -1)], axis=1)
    occ.sort(axis=1)
    pad = np.full((n, 1), -2, np.int32)
    prev = np.concatenate([pad, occ[:, :-1]], axis=1)
    nxt = np.concatenate([occ[:, 1:], pad], axis=1)
    single = (occ >= 0) & (occ != prev) & (occ != nxt)
    return ~single.any(axis=1)


def refine(records, body_predicates, max_atoms):
    """gpu_host_refine_rules: the canonical key records of every refinement
    of the given records (key_refinements, on the host's threads)."""
    from . import lib as L
    lib = L.load_library()
    P = np.ascontiguousarray(records, np.int32).reshape(-1, KEY_WORDS)
    B = np.ascontiguousarray(body_predicates, np.int32)
    cap = 4 * P.shape[0] * max(1, B.shape[0])   # a guess; a second call when it is short
    while True:
        out = np.empty((max(cap, 1), KEY_WORDS), np.int32)
        m = lib.gpu_host_refine_rules(P.ctypes.data, int(P.shape[0]), B.ctypes.data, int(B.shape[0]),
                                      int(max_atoms), out.ctypes.data, int(cap))
        if m < 0:
            raise RuntimeError("gpu_host_refine_rules failed")
        if m <= cap:
            return out[:m]
        cap = int(m)


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

    evaluate(records) -> structured array with fields status, support, headSize,
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

    def mine(self, evaluate):
        """evaluate(records) gets a level's key records (words_of_keys turns
        them into rule words)."""
        st = MiningStats()
        t0 = time.perf_counter()
        heads = np.zeros((len(self.heads), KEY_WORDS), np.int32)
        heads[:, 0] = self.heads
        level = refine(heads, self.bodies, self.max_atoms)
        out = []
        while len(level):
            t1 = time.perf_counter()
            res = evaluate(level)
            st.eval_s.append(time.perf_counter() - t1)
            st.candidates.append(int(level.shape[0]))
            ok = (res["status"] == 0) & (res["support"] >= self.min_support) & (res["headCoverage"] >= self.min_hc)
            report = ok & closed_mask(level) & (res["bodyPairs"] >= 0) & (res["confidence"] >= self.min_conf)
            for i in np.flatnonzero(report):
                r = res[i]
                out.append(MinedRule(rule_from_key(record_key(level[i])), int(r["support"]), int(r["headSize"]),
                                     int(r["bodySize"]), float(r["headCoverage"]), float(r["confidence"])))
            st.passed.append(int(ok.sum()))
            level = refine(level[ok], self.bodies, self.max_atoms) if ok.any() else level[:0]
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

    def __call__(self, records):
        words = words_of_keys(records, self.graph.n_entities)   # predicate ids follow the entities
        closed = closed_mask(records)
        out = np.zeros(len(records), self.L.RESULT64_DTYPE)
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

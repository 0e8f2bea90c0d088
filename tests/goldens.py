"""Loads the committed golden fixtures (tests/golden/*.json) and rebuilds the
graph and rules they were made from."""
from __future__ import annotations

import glob
import hashlib
import json
import os

import numpy as np

from masterthesis_gpu_acceleration_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Fixtures on the 6.8M-triple high-fan-out graph: the reference programs took
# up to minutes per rule on them, and so would the CPU oracle, so they are
# checked on the GPU only (against the reference programs' own values).
HIGH_FANOUT = ("hub_fanout_full", "serving_stream_400")


def names(cpu=False):
    """Fixture names; cpu=True leaves out the ones too large for the CPU oracle."""
    out = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*.json")))
    return [n for n in out if not (cpu and n in HIGH_FANOUT)]


_GEN = {}


def generate(fn, kwargs):
    """Generator output, once per process (the high-fan-out graph takes
    ~20 s to generate; the serving-stream fixture runs over it)."""
    key = (fn, json.dumps(kwargs, sort_keys=True))
    if key not in _GEN:
        kw = dict(kwargs)
        if fn == "serving_stream":
            kw["graph"] = generate("hub_fanout", {}).graph
        _GEN[key] = getattr(W, fn)(**kw)
    return _GEN[key]


def _sha(g):
    h = hashlib.sha256()
    for a in (g.s, g.p, g.o):
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


def rule_from_json(d):
    dec = lambda a: ((a[0][0], a[0][1] if a[0][0] == "v" else int(a[0][1])), int(a[1]),  # noqa: E731
                     (a[2][0], a[2][1] if a[2][0] == "v" else int(a[2][1])))
    return (dec(d["head"]), [dec(a) for a in d["body"]])


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        fx = json.load(f)
    if "triples" in fx:
        s, p, o = (np.asarray(x, np.int32) for x in fx["triples"])
        g = W.Graph(s, p, o, fx["n_entities"], fx["n_predicates"], name)
    else:
        gen = fx["generator"]
        g = generate(gen["fn"], gen["kwargs"]).graph
    assert _sha(g) == fx["graph_sha256"], "generator no longer reproduces fixture graph " + name
    rules = [rule_from_json(r) for r in fx["rules"]]
    return fx, g, rules


def head_pair_in_body(rule):
    """Both head terms are variables that occur in the body."""
    head, body = rule
    bv = {t[1] for a in body for t in (a[0], a[2]) if t[0] == "v"}
    return all(t[0] == "v" and t[1] in bv for t in (head[0], head[2]))


def expected(fx, rule_idx, rule):
    """Canonical expectation for one rule from the two reference programs.

    support: CPU_01's count (Rewritten agrees except its minSupport=1 scan,
    which can report 0 where the true count is 1).
    bodyPairs (ungated): CPU_01's bodySize when both head variables occur in
    the body; otherwise 0 (Rewritten's rule; CPU_01 counts (s,-1) pairs when
    only the head object is missing from the body).
    """
    c = fx["cpu01"][rule_idx]
    sup = c[0]
    pairs = c[3] if head_pair_in_body(rule) else 0
    return {"support": sup, "headSize": c[1], "headSupport": c[2], "bodyPairs": pairs,
            "bodySize": pairs if sup > 0 else 0}

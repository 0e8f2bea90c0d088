"""The allocation-free per-rule entry point (gpu_evaluate_rule_noalloc)
against gpu_evaluate_rule on the same Turtle-loaded graph: every field of
the 56-byte result (device times aside) is equal, call after call, also
when a call with many seeds is followed by one with fewer (the shared device
scratch keeps the last call's contents), and both equal the reference
programs' outputs. Exact integer counts and IEEE measures: tolerance 0.
A rule with more seeds than the entry point's limit is refused (-1) and
the context stays usable."""
import ctypes
import os
import sys
import tempfile

import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import workloads as W

import goldens

pytestmark = pytest.mark.gpu

FIELDS = ("support", "headSize", "headSupport", "headCoverage", "bodySize", "confidence")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fields(r):
    return tuple(getattr(r, f) for f in FIELDS)


def noalloc(api, ctx, text):
    out = L.GpuRuleResult()
    rc = api.lib.gpu_evaluate_rule_noalloc(ctx, text.encode(), ctypes.byref(out))
    return rc, out


@pytest.mark.parametrize("name", ["dbpedia_shaped_5k", "tiny_3", "small_paths_2k"])
def test_noalloc_matches_evaluate_rule(name):
    say("noalloc", name)
    fx, g, rules = goldens.load(name)
    api = L.GpuRulesLib()
    with tempfile.TemporaryDirectory() as d:
        ttl = os.path.join(d, "g.ttl")
        W.write_ttl(g, ttl)
        ctx = api.load_graph(ttl)
    try:
        texts = fx["rules_text"]
        # largest head first, then smaller ones: each call follows one with more seeds
        order = sorted(range(len(rules)), key=lambda i: -fx["cpu01"][i][1])
        for i in order:
            ref = api.evaluate_rule(ctx, texts[i])
            rc, got = noalloc(api, ctx, texts[i])
            assert rc == 0, (name, i)
            assert fields(got) == fields(ref), (name, i, fields(got), fields(ref), texts[i])
            exp = goldens.expected(fx, i, rules[i])
            assert (got.support, got.headSize, got.headSupport, got.bodySize) == (
                exp["support"], exp["headSize"], exp["headSupport"], exp["bodySize"]), (name, i)
        # two no-alloc calls in a row, the second with fewer seeds, then the first again
        big, small = order[0], order[-1]
        r1 = noalloc(api, ctx, texts[big])
        r2 = noalloc(api, ctx, texts[small])
        r3 = noalloc(api, ctx, texts[big])
        assert r1[0] == r2[0] == r3[0] == 0
        assert fields(r2[1]) == fields(api.evaluate_rule(ctx, texts[small]))
        assert fields(r1[1]) == fields(r3[1]) == fields(api.evaluate_rule(ctx, texts[big]))
        # malformed rule -> -1, as gpu_evaluate_rule
        assert noalloc(api, ctx, "( ?a <p> ?b")[0] == -1
    finally:
        api.free_graph(ctx)


def test_noalloc_refuses_beyond_its_limit():
    say("noalloc-limit")
    fx, g, rules = goldens.load("dbpedia_shaped_5k")
    api = L.GpuRulesLib()
    with tempfile.TemporaryDirectory() as d:
        ttl = os.path.join(d, "g.ttl")
        W.write_ttl(g, ttl)
        ctx = api.load_graph(ttl)
    try:
        # a rule with support ran a support launch over all its head triples (more than the limit)
        i = max((k for k in range(len(rules)) if fx["cpu01"][k][0] > 0), key=lambda k: fx["cpu01"][k][1])
        assert fx["cpu01"][i][1] > 8
        ref = api.evaluate_rule(ctx, fx["rules_text"][i])
        assert api.lib.gpu_set_tuning(ctx, b"noalloc_max_seeds", 8) == 0
        rc, out = noalloc(api, ctx, fx["rules_text"][i])
        assert rc == -1 and out.support == 0 and out.bodySize == 0
        assert api.lib.gpu_set_tuning(ctx, b"noalloc_max_seeds", ((16 << 20) - 512) // 8) == 0   # all it holds
        rc, out = noalloc(api, ctx, fx["rules_text"][i])
        assert rc == 0 and fields(out) == fields(ref)
        assert api.lib.gpu_set_tuning(ctx, b"noalloc_max_seeds", 1 << 40) == -1   # beyond the scratch
    finally:
        api.free_graph(ctx)

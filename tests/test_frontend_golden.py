"""The library's Turtle reader and rule parser against the reference's own
(GPU_03/RdfIndexes.cpp, GPU_03/RuleParser.cpp), through the golden output
in tests/golden/frontend/probe.json: every term id, the graph totals, and
every parsed rule (variable numbering, prefix expansion, interning of
unknown terms, parse errors)."""
import ctypes
import json
import os
import tempfile

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "frontend", "probe.json")
CASES = json.load(open(PROBE))


def ref_words(parsed):
    """Probe line [nbody, hs, hp, ho, (s p o)*] -> GPU_RULE_WORDS record."""
    w = np.zeros(L.RULE_WORDS, np.int32)
    enc = lambda t: -(int(t[1:]) + 1) if t[0] == "v" else int(t[1:])  # noqa: E731
    nb = int(parsed[0])
    w[0] = nb
    w[1:4] = [enc(t) for t in parsed[1:4]]
    for j in range(nb):
        w[4 + 3 * j:7 + 3 * j] = [enc(t) for t in parsed[4 + 3 * j:7 + 3 * j]]
    return w


@pytest.mark.parametrize("k", range(len(CASES)))
def test_frontend_matches_reference(k):
    case = CASES[k]
    lib = L.load_library()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.ttl")
        open(p, "w").write(case["ttl"])
        h = lib.gpu_host_load_turtle(p.encode())
        assert h
        try:
            for i, t in enumerate(case["terms"]):
                assert lib.gpu_host_term_id(h, t.encode()) == i, (i, t)
            nt, np_, ntr = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
            lib.gpu_host_graph_info(h, ctypes.byref(nt), ctypes.byref(np_), ctypes.byref(ntr))
            assert [nt.value, np_.value, ntr.value] == case["totals"]
            for rule, parsed in zip(case["rules"], case["parsed"]):
                w = np.zeros(L.RULE_WORDS, np.int32)
                rc = lib.gpu_host_parse_rule(h, rule.encode(), 0, w.ctypes.data)
                if parsed == ["error"]:
                    assert rc == -1, rule
                    continue
                assert rc == 0, rule
                assert list(w) == list(ref_words(parsed)), (rule, list(w), parsed)
        finally:
            lib.gpu_host_free(h)

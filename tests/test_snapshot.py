"""Binary graph snapshots (gpu_save_snapshot / gpu_load_snapshot /
gpu_host_load_snapshot), on the CPU: a snapshot of a Turtle-loaded or
term-id graph loads back to the same graph -- every term id, the graph
totals, every parsed rule, every compiled plan and the traffic model's bytes
per rule -- and damaged files are refused before anything is kept (the
checks that keep out-of-range ids away from the kernels)."""
import json
import os
import struct
import sys
import tempfile

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import goldens  # noqa: E402

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "frontend", "probe.json")
CASES = json.load(open(PROBE))

SECTIONS = ["predTerm", "predOff", "spoKey", "spoVal", "posKey", "posVal", "termOff", "termBytes", "prefixes"]


def header(raw):
    """Header fields of a snapshot: counts and section (offset, length)."""
    magic, version, bom = struct.unpack_from("<8sII", raw, 0)
    nE, T, P, nTerms, nPrefixes = struct.unpack_from("<5q", raw, 16)
    k = len(SECTIONS)
    offs = struct.unpack_from("<%dq" % k, raw, 56)
    lens = struct.unpack_from("<%dq" % k, raw, 56 + 8 * k)
    return dict(magic=magic, version=version, nE=nE, T=T, P=P, nTerms=nTerms,
                sec={n: (o, ln) for n, o, ln in zip(SECTIONS, offs, lens)})


def section(raw, h, name, dtype=np.int32):
    o, ln = h["sec"][name]
    return np.frombuffer(raw, dtype, ln // np.dtype(dtype).itemsize, o)


def plan_info(hg, words):
    out = np.zeros((words.shape[0], 7), np.int64)
    for i in range(words.shape[0]):
        w = np.ascontiguousarray(words[i], np.int32)
        assert hg.lib.gpu_host_plan_info(hg.h, w.ctypes.data, out[i].ctypes.data) == 0
    return out


@pytest.mark.parametrize("k", range(len(CASES)))
def test_turtle_snapshot_round_trip(k):
    """Term ids, totals and parsed rules (prefixes, literals, blank nodes) of
    every front-end probe graph are the same after a snapshot round trip."""
    case = CASES[k]
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.ttl")
        open(p, "w").write(case["ttl"])
        a = L.HostGraph.from_turtle(p)
        snap = os.path.join(d, "g.snap")
        a.save_snapshot(snap)
        b = L.HostGraph.from_snapshot(snap)
        try:
            assert a.info() == b.info() == tuple(case["totals"])
            for i, t in enumerate(case["terms"]):
                assert b.lib.gpu_host_term_id(b.h, t.encode()) == i, (i, t)
            for rule in case["rules"]:
                wa = np.zeros(L.RULE_WORDS, np.int32)
                wb = np.zeros(L.RULE_WORDS, np.int32)
                ra = a.lib.gpu_host_parse_rule(a.h, rule.encode(), 0, wa.ctypes.data)
                rb = b.lib.gpu_host_parse_rule(b.h, rule.encode(), 0, wb.ctypes.data)
                assert ra == rb and list(wa) == list(wb), rule
        finally:
            a.close()
            b.close()


@pytest.mark.parametrize("name", ["tiny_0", "tiny_3", "small_paths_2k", "dense_conf_300", "dbpedia_shaped_5k"])
def test_term_id_snapshot_round_trip(name):
    """Term-id graphs: same totals, same compiled plans (head figures, seed
    counts) and the same traffic-model bytes for every golden rule."""
    fx, g, rules = goldens.load(name)
    words = L.rule_words_for_graph(rules, g)
    a = L.HostGraph(g)
    with tempfile.TemporaryDirectory() as d:
        snap = os.path.join(d, "g.snap")
        a.save_snapshot(snap)
        raw = open(snap, "rb").read()
        h = header(raw)
        assert h["magic"] == b"GRSNAP1\0" and h["T"] == a.info()[2]
        b = L.HostGraph.from_snapshot(snap)
        try:
            assert a.info() == b.info()
            assert np.array_equal(plan_info(a, words), plan_info(b, words))
            assert np.array_equal(a.traffic(words), b.traffic(words))
        finally:
            a.close()
            b.close()


def test_snapshot_arrays_are_the_sorted_graph():
    """Per predicate, the SPO section holds the distinct (s, o) of its
    triples sorted by (s, o), and the POS section the distinct (o, s) sorted
    by (o, s) -- checked against numpy.unique of the raw triples."""
    fx, g, rules = goldens.load("small_paths_2k")
    a = L.HostGraph(g)
    with tempfile.TemporaryDirectory() as d:
        snap = os.path.join(d, "g.snap")
        a.save_snapshot(snap)
        raw = open(snap, "rb").read()
    a.close()
    h = header(raw)
    pt = section(raw, h, "predTerm")
    off = section(raw, h, "predOff")
    sk, sv = section(raw, h, "spoKey"), section(raw, h, "spoVal")
    pk, pv = section(raw, h, "posKey"), section(raw, h, "posVal")
    assert sorted((pt - g.n_entities).tolist()) == sorted(np.unique(g.p).tolist())
    for p in range(len(pt)):
        m = g.p == pt[p] - g.n_entities
        r = slice(off[p], off[p + 1])
        so = np.unique(np.stack([g.s[m], g.o[m]], 1), axis=0)
        os_ = np.unique(np.stack([g.o[m], g.s[m]], 1), axis=0)
        assert np.array_equal(np.stack([sk[r], sv[r]], 1), so), p
        assert np.array_equal(np.stack([pk[r], pv[r]], 1), os_), p
    assert off[-1] == len(np.unique(np.stack([g.p, g.s, g.o], 1), axis=0))


def _damaged(raw, how):
    h = header(raw)
    b = bytearray(raw)
    if how == "truncated":
        return bytes(b[: len(b) // 2])
    if how == "magic":
        b[0:8] = b"NOTSNAP\0"
    elif how == "id_out_of_range":
        o, _ = h["sec"]["spoVal"]
        struct.pack_into("<i", b, o, h["nE"] + 5)
    elif how == "negative_id":
        o, _ = h["sec"]["posKey"]
        struct.pack_into("<i", b, o, -1)
    elif how == "unsorted":
        # swap two neighbouring keys of one predicate: ids stay in range,
        # the row order breaks
        off = np.frombuffer(raw, np.int32, h["P"] + 1, h["sec"]["predOff"][0])
        key = np.frombuffer(raw, np.int32, h["T"], h["sec"]["spoKey"][0])
        inner = set(range(h["T"] - 1)) - set((off[1:] - 1).tolist())
        e = next(e for e in sorted(inner) if key[e] < key[e + 1])
        struct.pack_into("<ii", b, h["sec"]["spoKey"][0] + 4 * e, int(key[e + 1]), int(key[e]))
    elif how == "not_transposed":
        # change one POS value without breaking its row's order: the two
        # directions then hold different triples
        off = np.frombuffer(raw, np.int32, h["P"] + 1, h["sec"]["predOff"][0])
        key = np.frombuffer(raw, np.int32, h["T"], h["sec"]["posKey"][0])
        val = np.frombuffer(raw, np.int32, h["T"], h["sec"]["posVal"][0])
        for e in range(h["T"] - 1):
            if key[e] == key[e + 1] and val[e] + 1 < val[e + 1]:
                struct.pack_into("<i", b, h["sec"]["posVal"][0] + 4 * e, int(val[e]) + 1)
                break
        else:
            raise AssertionError("no entry to change")
    elif how == "section_past_end":
        struct.pack_into("<q", b, 56 + 8 * SECTIONS.index("posVal"), len(b) + 64)
    return bytes(b)


@pytest.mark.parametrize("how", ["truncated", "magic", "id_out_of_range", "negative_id", "unsorted",
                                 "not_transposed", "section_past_end"])
def test_damaged_snapshot_is_refused(how):
    fx, g, rules = goldens.load("small_paths_2k")
    a = L.HostGraph(g)
    lib = a.lib
    with tempfile.TemporaryDirectory() as d:
        snap = os.path.join(d, "g.snap")
        a.save_snapshot(snap)
        a.close()
        raw = open(snap, "rb").read()
        ok = lib.gpu_host_load_snapshot(snap.encode())   # the intact file loads
        assert ok
        lib.gpu_host_free(ok)
        bad = os.path.join(d, "bad.snap")
        open(bad, "wb").write(_damaged(raw, how))
        assert not lib.gpu_host_load_snapshot(bad.encode()), how
        assert not lib.gpu_host_load_snapshot(os.path.join(d, "missing.snap").encode())
        assert lib.gpu_save_snapshot(None, snap.encode()) == -1

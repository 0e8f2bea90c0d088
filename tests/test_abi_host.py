"""C-ABI checks that need no GPU: the library loads and exports every
function include/gpu_rules_api.h declares, the result struct keeps the
reference's 56-byte layout (checked from C with static_assert and from
ctypes), and the host side -- Turtle reader, term ids, rule parsers, head
figures -- behaves as the reference's front end does."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import workloads as W
from oracle import Oracle

import goldens

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpu_rules_api.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gpu_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = L.load_library()
    names = declared_functions()
    assert len(names) >= 5
    for n in ("gpu_load_graph", "gpu_evaluate_rule", "gpu_evaluate_rule_json", "gpu_free_graph", "gpu_query_vram"):
        assert n in names
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    assert set(names) <= exported
    assert set(L.exported_names()) <= exported


def test_count_decisions_with_forced_counters():
    """The host's decisions on a pass's device counters (taken by the batch,
    per-rule and serving paths after every pass), fed counter values no test
    graph reaches: key counts past 2^32 must never be taken as counted, and
    past the largest capacity they must fail rather than regrow."""
    act = L.load_library().gpu_host_count_action
    cap_max = 1 << 31
    SORT, HEAVY = 0, 1
    OK, REGROW, FAIL = 0, 1, 2
    assert act(SORT, 0, 4096, cap_max) == OK
    assert act(SORT, 4096, 4096, cap_max) == OK
    assert act(SORT, 4097, 4096, cap_max) == REGROW
    assert act(SORT, cap_max, cap_max, cap_max) == OK
    assert act(SORT, cap_max + 1, cap_max, cap_max) == FAIL
    # a count that a 32-bit counter would have wrapped to a small value
    for emitted in ((1 << 32) + 5, 5_370_000_000, (1 << 40)):
        assert act(SORT, emitted, 1 << 22, cap_max) == REGROW
        assert act(SORT, emitted, cap_max, cap_max) == FAIL
    # a lowered maximum (the "sort_cap_max" tuning key) fails at that capacity
    assert act(SORT, 5000, 4096, 4096) == FAIL
    assert act(HEAVY, 10, 10, 0) == OK
    assert act(HEAVY, 11, 10, 0) == FAIL
    assert act(HEAVY, (1 << 32) + 1, 0xFFFFFFF0, 0) == FAIL
    assert act(2, 0, 0, 0) == -1
    assert act(SORT, -1, 0, 0) == -1


def test_result_struct_layout_ctypes():
    assert ctypes.sizeof(L.GpuRuleResult) == 56
    offs = [getattr(L.GpuRuleResult, f).offset for f, _ in L.GpuRuleResult._fields_]
    assert offs == [0, 4, 8, 16, 24, 32, 40, 48]


def test_header_compiles_as_c_and_links():
    """The header is self-contained plain C; a C caller links against the
    library (no GPU call beyond gpu_query_vram, which fails cleanly here)."""
    prog = r'''
#include "gpu_rules_api.h"
#include <stdio.h>
#include <stddef.h>
_Static_assert(sizeof(GpuRuleResult) == 56, "size");
_Static_assert(offsetof(GpuRuleResult, support) == 0, "o0");
_Static_assert(offsetof(GpuRuleResult, headSize) == 4, "o1");
_Static_assert(offsetof(GpuRuleResult, headSupport) == 8, "o2");
_Static_assert(offsetof(GpuRuleResult, headCoverage) == 16, "o3");
_Static_assert(offsetof(GpuRuleResult, bodySize) == 24, "o4");
_Static_assert(offsetof(GpuRuleResult, confidence) == 32, "o5");
_Static_assert(offsetof(GpuRuleResult, supportMs) == 40, "o6");
_Static_assert(offsetof(GpuRuleResult, confidenceMs) == 48, "o7");
_Static_assert(sizeof(GpuRuleDesc) == 4 * GPU_RULE_WORDS, "desc");
int main(void) {
    size_t f = 0, t = 0;
    int rc = gpu_query_vram(&f, &t);
    printf("%s rc=%d\n", gpu_rules_version(), rc);
    return (gpu_load_graph(NULL) == NULL) ? 0 : 1;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(prog)
        subprocess.check_call(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), c, "-o", exe,
                               L.LIB_PATH, "-Wl,-rpath," + os.path.dirname(L.LIB_PATH)])
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "gpu_rules" in r.stdout


TTL = """@prefix ex: <https://ex.org/> .
PREFIX rdf: <http://www.w3.org/1999/02/22-rdf-syntax-ns#>
# comment line
ex:a ex:knows ex:b , ex:c ;
     a ex:Person .
<https://ex.org/b> ex:age 42 ; ex:name "Bob"@en .
_:n1 ex:knows [] .
ex:c ex:label "x y"^^<http://www.w3.org/2001/XMLSchema#string> ; ex:flag true .
ex:a ex:knows ex:b .
"""


def host_graph(text):
    lib = L.load_library()
    d = tempfile.mkdtemp()
    p = os.path.join(d, "g.ttl")
    open(p, "w").write(text)
    h = lib.gpu_host_load_turtle(p.encode())
    assert h
    return lib, h


def test_turtle_ids_in_first_seen_order():
    """ids follow subject, predicate, object of each triple in file order
    (GPU_03/RdfIndexes.cpp:172-178); term strings follow the reference
    reader's conventions (IRIs unbracketed, prefixes expanded, `a`, quoted
    literals with @lang / ^^type, blank nodes, numbers, booleans)."""
    lib, h = host_graph(TTL)
    expect = ["https://ex.org/a", "https://ex.org/knows", "https://ex.org/b", "https://ex.org/c",
              "http://www.w3.org/1999/02/22-rdf-syntax-ns#type", "https://ex.org/Person",
              "https://ex.org/age", '"42"', "https://ex.org/name", '"Bob"@en', "_:n1", "_:anon1",
              "https://ex.org/label", '"x y"^^http://www.w3.org/2001/XMLSchema#string', "https://ex.org/flag",
              '"true"^^http://www.w3.org/2001/XMLSchema#boolean']
    for i, t in enumerate(expect):
        assert lib.gpu_host_term_id(h, t.encode()) == i, t
    nt, np_, ntr = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
    lib.gpu_host_graph_info(h, ctypes.byref(nt), ctypes.byref(np_), ctypes.byref(ntr))
    assert nt.value == len(expect)
    assert np_.value == 6          # knows, type, age, name, label, flag
    assert ntr.value == 8          # 9 statements, one duplicate
    lib.gpu_host_free(h)


def parse(lib, h, text, is_json=False):
    w = np.zeros(L.RULE_WORDS, np.int32)
    rc = lib.gpu_host_parse_rule(h, text.encode(), int(is_json), w.ctypes.data)
    return rc, w


def test_rule_parser_numbering_and_prefixes():
    lib, h = host_graph(TTL)
    ida = lib.gpu_host_term_id(h, b"https://ex.org/a")
    knows = lib.gpu_host_term_id(h, b"https://ex.org/knows")
    # variables numbered body-first, subject before object; prefixes from the TTL
    rc, w = parse(lib, h, "( ?x ex:knows ?z ) ^ ( ?z <https://ex.org/knows> ?y ) => ( ?x ex:knows ?y ) | note")
    assert rc == 0
    assert list(w[:10]) == [2, -1, knows, -3, -1, knows, -2, -2, knows, -3]
    rc, w = parse(lib, h, "( ex:a ex:knows ?b ) => ( ?b ex:knows ex:a )")
    assert rc == 0 and w[4] == ida and w[6] == -1 and w[1] == -1 and w[3] == ida
    # JSON in the RDFRules export shape
    js = ('{"body":[{"subject":{"type":"variable","value":"?x"},"predicate":"<https://ex.org/knows>",'
          '"object":{"type":"variable","value":"?y"}}],"head":{"subject":{"type":"variable","value":"?y"},'
          '"predicate":"ex:knows","object":{"type":"constant","value":"<https://ex.org/a>"}}}')
    rc, w = parse(lib, h, js, True)
    assert rc == 0 and list(w[:7]) == [1, -2, knows, ida, -1, knows, -2]
    # malformed
    assert parse(lib, h, "( ?a ex:knows ?b ) ( ?b")[0] == -1
    assert parse(lib, h, "no arrow here")[0] == -1
    # unknown constants get fresh ids past the graph's terms
    rc, w = parse(lib, h, "( ?a <https://ex.org/unknown> ?b ) => ( ?a ex:knows ?b )")
    assert rc == 0 and w[5] >= 16
    lib.gpu_host_free(h)


@pytest.mark.parametrize("name", [n for n in goldens.names() if n.startswith("tiny")][:3] + ["small_paths_2k"])
def test_head_figures_and_skip_rules(name):
    """headSize / headSupport come from the host planner; check them against
    the reference programs' outputs through a Turtle-loaded host graph."""
    fx, g, rules = goldens.load(name)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.ttl")
        W.write_ttl(g, p)
        lib = L.load_library()
        h = lib.gpu_host_load_turtle(p.encode())
        info = np.zeros(8, np.int64)
        for i, text in enumerate(fx["rules_text"]):
            rc, w = parse(lib, h, text)
            assert rc == 0
            lib.gpu_host_plan_info(h, w.ctypes.data, info.ctypes.data)
            c = fx["cpu01"][i]
            assert (info[0], info[1]) == (c[1], c[2]), (name, i, text)
            assert info[2] == 0
            if c[0] > 0:
                assert info[3] == 1   # a rule with support is never statically zero
        # 9 body atoms -> unsupported shape (status 1); 0 atoms likewise
        nine = " ^ ".join(["( ?a <https://ex.org/p/0> ?b )"] * 9) + " => ( ?a <https://ex.org/p/1> ?b )"
        rc, w = parse(lib, h, nine)
        assert rc == -1       # more than 8 atoms cannot be expressed as a word record
        words = np.zeros(L.RULE_WORDS, np.int32)
        words[0] = 0
        words[1:4] = (-1, lib.gpu_host_term_id(h, b"https://ex.org/p/1"), -2)
        lib.gpu_host_plan_info(h, words.ctypes.data, info.ctypes.data)
        assert info[2] == 1
        lib.gpu_host_free(h)


def test_no_gpu_context_fails_loudly():
    """Without a device the library refuses to build a context instead of
    falling back to any CPU path."""
    rc, _, _ = L.GpuRulesLib().query_vram()
    if rc == 0:
        pytest.skip("a GPU is present")
    g = W.tiny_random(1)
    with pytest.raises(RuntimeError):
        L.Engine.from_graph(g)


def test_counter_cli_usage_and_missing_graph():
    """gpu_rules_counter (the CPU_01 wire-format counter over the library)
    exits 1 with a message on missing arguments or an unreadable graph, as
    the reference program ends on an exception."""
    cli = os.path.join(ROOT, "masterthesis_gpu_acceleration_amd", "gpu_rules_counter")
    assert os.path.exists(cli)
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    with tempfile.TemporaryDirectory() as d:
        rf = os.path.join(d, "r.txt")
        open(rf, "w").write("( ?a <https://ex.org/p/0> ?b ) => ( ?a <https://ex.org/p/1> ?b )\n")
        r = subprocess.run([cli, os.path.join(d, "missing.ttl"), rf, "--json"], capture_output=True, text=True)
        assert r.returncode == 1 and r.stdout == ""


SCALA = os.path.join(ROOT, "masterthesis_gpu_acceleration_amd", "bindings", "scala", "GpuRulesBatchLib.scala")


def c_prototypes():
    """name -> list of C parameter declarations, from include/gpu_rules_api.h."""
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(gpu_\w+)\s*\(([^)]*)\)\s*;", src):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = params
    return out


def test_scala_binding_matches_header():
    """The Scala JNA binding of the additive entry points (not compiled here:
    no JVM) declares functions the header declares, with as many arguments,
    pointer arguments where the C parameter is a pointer and Int/Long where
    it is a 32/64-bit integer; its structures list the C fields in C order,
    as the ctypes binding does (whose sizes and offsets are pinned above)."""
    text = open(SCALA).read()
    protos = c_prototypes()
    decls = re.findall(r"def (gpu_\w+)\(([^)]*)\):\s*(\w+)", text, flags=re.S)
    assert len(decls) >= 15
    for name, args, ret in decls:
        assert name in protos, name
        params = protos[name]
        sargs = [a.strip() for a in args.replace("\n", " ").split(",") if a.strip()]
        assert len(sargs) == len(params), (name, sargs, params)
        for sa, cp in zip(sargs, params):
            stype = sa.split(":", 1)[1].strip()
            if "*" in cp:
                assert stype in ("Pointer", "String", "GpuRuleResultJ", "GpuRuleResult64", "GpuServeCompletion",
                                 "DoubleByReference", "LongByReference", "IntByReference", "Array[String]",
                                 "Array[Int]", "Array[Long]"), (name, sa, cp)
            elif cp.startswith("int64_t"):
                assert stype == "Long", (name, sa, cp)
            elif cp.startswith("int32_t") or cp.startswith("int "):
                assert stype == "Int", (name, sa, cp)
    def fields(cls):
        m = re.search(r"class %s extends Structure \{(.*?)override def getFieldOrder.*?asList\((.*?)\)" % cls, text,
                      flags=re.S)
        assert m, cls
        declared = re.findall(r"var (\w+):", m.group(1))
        order = re.findall(r'"(\w+)"', m.group(2))
        assert declared == order, cls
        return order
    assert fields("GpuRuleResultJ") == [f for f, _ in L.GpuRuleResult._fields_]
    assert fields("GpuRuleResult64") == [f for f, _ in L.GpuRuleResult64._fields_]
    from masterthesis_gpu_acceleration_amd import serving as S
    assert fields("GpuServeCompletion") == list(S.COMPLETION_DTYPE.names)

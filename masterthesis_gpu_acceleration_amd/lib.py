"""ctypes binding of libgpu_rules.so.

`GpuRulesLib` mirrors the reference's JNA binding (GPU_03/GpuRulesLib.scala:
23-58: the GpuRuleResult structure in the same field order, and the
gpu_load_graph / gpu_evaluate_rule / gpu_free_graph / gpu_query_vram calls),
with the same C calling convention JNA uses, plus the library's additive
entry points. `Engine` is a small convenience layer over it for term-id
graphs and structured rules.

There is no fallback: if the library is missing or no GPU is present, the
calls fail loudly (RuntimeError / NULL context).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GPU_RULES_LIB", os.path.join(PKG, "libgpu_rules.so"))

RULE_WORDS = 28
UNGATED = 1
PER_RULE = 2


class GpuRuleResult(ctypes.Structure):
    """GPU_03/gpu_rules_api.h:21-30 (56 bytes)."""
    _fields_ = [("support", ctypes.c_int), ("headSize", ctypes.c_int), ("headSupport", ctypes.c_int),
                ("headCoverage", ctypes.c_double), ("bodySize", ctypes.c_int), ("confidence", ctypes.c_double),
                ("supportMs", ctypes.c_double), ("confidenceMs", ctypes.c_double)]


class GpuRuleResult64(ctypes.Structure):
    _fields_ = [("support", ctypes.c_int64), ("headSize", ctypes.c_int64), ("headSupport", ctypes.c_int64),
                ("bodySize", ctypes.c_int64), ("bodyPairs", ctypes.c_int64), ("headCoverage", ctypes.c_double),
                ("confidence", ctypes.c_double), ("status", ctypes.c_int32), ("latencyUs", ctypes.c_int32)]


RESULT64_DTYPE = np.dtype([("support", "<i8"), ("headSize", "<i8"), ("headSupport", "<i8"), ("bodySize", "<i8"),
                           ("bodyPairs", "<i8"), ("headCoverage", "<f8"), ("confidence", "<f8"),
                           ("status", "<i4"), ("latencyUs", "<i4")])
assert RESULT64_DTYPE.itemsize == ctypes.sizeof(GpuRuleResult64)

# name -> (restype, argtypes)
_SIGS = {
    "gpu_load_graph": (ctypes.c_void_p, [ctypes.c_char_p]),
    "gpu_evaluate_rule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(GpuRuleResult)]),
    "gpu_evaluate_rule_json": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(GpuRuleResult)]),
    "gpu_evaluate_rule_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(GpuRuleResult)]),
    "gpu_free_graph": (None, [ctypes.c_void_p]),
    "gpu_query_vram": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "gpu_load_graph_triples": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                 ctypes.c_int64]),
    "gpu_evaluate_rules": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_void_p]),
    "gpu_evaluate_rules_text": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int32,
                                               ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]),
    "gpu_last_batch_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_int64)]),
    "gpu_batch_create": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "gpu_evaluate_rules_min": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_int64, ctypes.c_double, ctypes.c_void_p]),
    "gpu_batch_create_min": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int64, ctypes.c_double]),
    "gpu_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_batch_free": (None, [ctypes.c_void_p]),
    "gpu_copy_probe": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_int64]),
    "gpu_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                    ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "gpu_last_kernel_ms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_last_rule_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "gpu_last_dedup_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "gpu_plan_traffic": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "gpu_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    "gpu_get_tuning": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_char_p]),
    "gpu_host_load_turtle": (ctypes.c_void_p, [ctypes.c_char_p]),
    "gpu_host_load_triples": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                                ctypes.c_int64]),
    "gpu_host_load_snapshot": (ctypes.c_void_p, [ctypes.c_char_p]),
    "gpu_save_snapshot": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "gpu_load_snapshot": (ctypes.c_void_p, [ctypes.c_char_p]),
    "gpu_host_free": (None, [ctypes.c_void_p]),
    "gpu_host_graph_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]),
    "gpu_host_term_id": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_char_p]),
    "gpu_host_parse_rule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p]),
    "gpu_host_plan_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_host_rules_work": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "gpu_host_count_action": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
    "gpu_host_support_threshold": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, ctypes.c_double]),
    "gpu_host_body_groundings": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "gpu_host_refine_rules": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                               ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64]),
    "gpu_host_term_string": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]),
    "gpu_rules_version": (ctypes.c_char_p, []),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libgpu_rules.so and declares every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libgpu_rules.so is not built (%s); run masterthesis_gpu_acceleration_amd.build" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_names():
    return list(_SIGS)


class GpuRulesLib:
    """The reference binding's surface (GpuRulesLib.scala:46-116): load a
    graph once, evaluate rules, free. evaluate_rule raises on a non-zero
    return code, as the Scala wrapper does (:107-112)."""

    def __init__(self, path: str = LIB_PATH):
        self.lib = load_library(path)

    def query_vram(self):
        f = ctypes.c_size_t()
        t = ctypes.c_size_t()
        rc = self.lib.gpu_query_vram(ctypes.byref(f), ctypes.byref(t))
        return rc, f.value, t.value

    def load_graph(self, ttl_path: str):
        ctx = self.lib.gpu_load_graph(ttl_path.encode())
        if not ctx:
            raise RuntimeError("gpu_load_graph failed for " + ttl_path)
        return ctx

    def evaluate_rule(self, ctx, rule_text: str) -> GpuRuleResult:
        out = GpuRuleResult()
        rc = self.lib.gpu_evaluate_rule(ctx, rule_text.encode(), ctypes.byref(out))
        if rc != 0:
            raise RuntimeError("gpu_evaluate_rule returned %d" % rc)
        return out

    def evaluate_rule_json(self, ctx, rule_json: str) -> GpuRuleResult:
        out = GpuRuleResult()
        rc = self.lib.gpu_evaluate_rule_json(ctx, rule_json.encode(), ctypes.byref(out))
        if rc != 0:
            raise RuntimeError("gpu_evaluate_rule_json returned %d" % rc)
        return out

    def free_graph(self, ctx):
        self.lib.gpu_free_graph(ctx)


def _body_groundings(lib, handle, rule_words):
    w = np.ascontiguousarray(rule_words, np.int32)
    out = np.zeros(w.shape[0], np.int64)
    if lib.gpu_host_body_groundings(handle, w.ctypes.data, int(w.shape[0]), out.ctypes.data) != 0:
        raise RuntimeError("gpu_host_body_groundings failed")
    return out


class HostGraph:
    """Host-only graph handle (no device): planner and traffic model."""

    def __init__(self, g):
        self.lib = load_library()
        s = np.ascontiguousarray(g.s, np.int32)
        p = np.ascontiguousarray(g.p.astype(np.int64) + g.n_entities, np.int32)
        o = np.ascontiguousarray(g.o, np.int32)
        self.h = self.lib.gpu_host_load_triples(s.ctypes.data, p.ctypes.data, o.ctypes.data, int(s.shape[0]),
                                                int(g.n_entities + g.n_predicates))
        if not self.h:
            raise RuntimeError("gpu_host_load_triples failed")

    @classmethod
    def _wrap(cls, h, what: str):
        if not h:
            raise RuntimeError(what + " failed")
        obj = cls.__new__(cls)
        obj.lib = load_library()
        obj.h = h
        return obj

    @classmethod
    def from_turtle(cls, path: str):
        return cls._wrap(load_library().gpu_host_load_turtle(path.encode()), "gpu_host_load_turtle")

    @classmethod
    def from_snapshot(cls, path: str):
        return cls._wrap(load_library().gpu_host_load_snapshot(path.encode()), "gpu_host_load_snapshot")

    def save_snapshot(self, path: str):
        if self.lib.gpu_save_snapshot(self.h, path.encode()) != 0:
            raise RuntimeError("gpu_save_snapshot failed for " + path)

    def info(self):
        """(terms, predicates, triples)."""
        nt, np_, ntr = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
        if self.lib.gpu_host_graph_info(self.h, ctypes.byref(nt), ctypes.byref(np_), ctypes.byref(ntr)) != 0:
            raise RuntimeError("gpu_host_graph_info failed")
        return nt.value, np_.value, ntr.value

    def traffic(self, rule_words: np.ndarray) -> np.ndarray:
        """Per rule: [support bytes, pair bytes, support nodes, pair nodes]."""
        w = np.ascontiguousarray(rule_words, np.int32)
        out = np.zeros((w.shape[0], 4), np.int64)
        if self.lib.gpu_plan_traffic(self.h, w.ctypes.data, int(w.shape[0]), out.ctypes.data) != 0:
            raise RuntimeError("gpu_plan_traffic failed")
        return out

    def body_groundings(self, rule_words: np.ndarray) -> np.ndarray:
        """Per rule: injective body groundings -- the pairs the reference's
        body-size kernel emits before dedup (-1 where not defined)."""
        return _body_groundings(self.lib, self.h, rule_words)

    def work(self, rule_words: np.ndarray) -> np.ndarray:
        """Per rule: estimated candidates its searches enumerate (the
        figure the server schedules by)."""
        w = np.ascontiguousarray(rule_words, np.int32)
        out = np.zeros(w.shape[0], np.float64)
        if self.lib.gpu_host_rules_work(self.h, w.ctypes.data, int(w.shape[0]), out.ctypes.data) != 0:
            raise RuntimeError("gpu_host_rules_work failed")
        return out

    def close(self):
        if self.h:
            self.lib.gpu_host_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """A resident graph plus batched / per-rule evaluation of structured
    rules (int32 records of RULE_WORDS words, see workloads.encode_rules)."""

    def __init__(self, ctx, lib=None):
        self.lib = lib or load_library()
        self.ctx = ctx

    @classmethod
    def from_ttl(cls, path: str):
        lib = load_library()
        ctx = lib.gpu_load_graph(path.encode())
        if not ctx:
            raise RuntimeError("gpu_load_graph failed for " + path)
        return cls(ctx, lib)

    @classmethod
    def from_snapshot(cls, path: str):
        lib = load_library()
        ctx = lib.gpu_load_snapshot(path.encode())
        if not ctx:
            raise RuntimeError("gpu_load_snapshot failed for " + path)
        return cls(ctx, lib)

    def save_snapshot(self, path: str):
        if self.lib.gpu_save_snapshot(self.ctx, path.encode()) != 0:
            raise RuntimeError("gpu_save_snapshot failed for " + path)

    @classmethod
    def from_triples(cls, s, p, o, n_terms: int):
        lib = load_library()
        s = np.ascontiguousarray(s, np.int32)
        p = np.ascontiguousarray(p, np.int32)
        o = np.ascontiguousarray(o, np.int32)
        ctx = lib.gpu_load_graph_triples(s.ctypes.data, p.ctypes.data, o.ctypes.data, int(s.shape[0]), int(n_terms))
        if not ctx:
            raise RuntimeError("gpu_load_graph_triples failed (no GPU, HBM short, or bad ids)")
        return cls(ctx, lib)

    @classmethod
    def from_graph(cls, g):
        """workloads.Graph: entity i -> term i, predicate k -> term n_entities + k."""
        return cls.from_triples(g.s, g.p.astype(np.int64) + g.n_entities, g.o, g.n_entities + g.n_predicates)

    def evaluate(self, rule_words: np.ndarray, flags: int = 0) -> np.ndarray:
        w = np.ascontiguousarray(rule_words, np.int32)
        n = int(w.shape[0])
        out = np.zeros(n, RESULT64_DTYPE)
        rc = self.lib.gpu_evaluate_rules(self.ctx, w.ctypes.data, n, int(flags), out.ctypes.data)
        if rc != 0:
            raise RuntimeError("gpu_evaluate_rules returned %d" % rc)
        return out

    def evaluate_min(self, rule_words: np.ndarray, min_support: int = 1, min_head_coverage: float = 0.0,
                     flags: int = 0) -> np.ndarray:
        """gpu_evaluate_rules_min: exact support of every rule, body pairs only
        for rules with support >= min_support and head coverage >=
        min_head_coverage (bodyPairs -1 for the others)."""
        w = np.ascontiguousarray(rule_words, np.int32)
        n = int(w.shape[0])
        out = np.zeros(n, RESULT64_DTYPE)
        rc = self.lib.gpu_evaluate_rules_min(self.ctx, w.ctypes.data, n, int(flags), int(min_support),
                                             float(min_head_coverage), out.ctypes.data)
        if rc != 0:
            raise RuntimeError("gpu_evaluate_rules_min returned %d" % rc)
        return out

    def body_groundings(self, rule_words: np.ndarray) -> np.ndarray:
        """gpu_host_body_groundings on this context's graph (host threads)."""
        return _body_groundings(self.lib, self.ctx, rule_words)

    def evaluate_text(self, rules, is_json: bool = False, flags: int = 0) -> np.ndarray:
        n = len(rules)
        arr = (ctypes.c_char_p * n)(*[r.encode() for r in rules])
        out = np.zeros(n, RESULT64_DTYPE)
        rc = self.lib.gpu_evaluate_rules_text(self.ctx, arr, n, int(is_json), int(flags), out.ctypes.data)
        if rc != 0:
            raise RuntimeError("gpu_evaluate_rules_text returned %d" % rc)
        return out

    def batch(self, rule_words: np.ndarray, flags: int = 0, min_support=None,
              min_head_coverage: float = 0.0) -> "ResidentBatch":
        return ResidentBatch(self, rule_words, flags, min_support, min_head_coverage)

    def last_stats(self):
        t, s, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        h = ctypes.c_int64()
        self.lib.gpu_last_batch_stats(self.ctx, ctypes.byref(t), ctypes.byref(s), ctypes.byref(p), ctypes.byref(h))
        return {"total_ms": t.value, "support_ms": s.value, "pairs_ms": p.value, "heavy_seeds": h.value}

    def dedup_stats(self):
        """Sort-unique dedup of the last evaluation: radix sort + unique time
        (eager batches), keys emitted, key capacity of the buffers."""
        ms, n, cap = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        self.lib.gpu_last_dedup_stats(self.ctx, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(cap))
        return {"sort_ms": ms.value, "keys_emitted": n.value, "sort_cap": cap.value}

    def autotune(self, rule_words: np.ndarray, flags: int = 0):
        """Runs the library's autotuner on these rules; returns its trial log
        as rows of (budget, thread waves/CU, wave waves/CU, ms, escalated,
        pairs dedup mode, giant_after)."""
        w = np.ascontiguousarray(rule_words, np.int32)
        log = np.zeros((64, 7), np.float64)
        n = ctypes.c_int32(0)
        if self.lib.gpu_autotune(self.ctx, w.ctypes.data, int(w.shape[0]), int(flags), log.ctypes.data, 64,
                                 ctypes.byref(n)) != 0:
            raise RuntimeError("gpu_autotune failed")
        return log[: n.value]

    def rule_times(self, n: int):
        """(support ms, confidence ms) per rule of the last per-rule evaluation."""
        a = np.zeros(n, np.float64)
        b = np.zeros(n, np.float64)
        k = self.lib.gpu_last_rule_times(self.ctx, a.ctypes.data, b.ctypes.data, int(n))
        if k < 0:
            raise RuntimeError("gpu_last_rule_times failed")
        return a[:k], b[:k]

    def kernel_ms(self):
        """[support thread, support wave, pairs thread, pairs wave] of the last
        eager batch (ms)."""
        out = np.zeros(4, np.float64)
        self.lib.gpu_last_kernel_ms(self.ctx, out.ctypes.data)
        return out

    def traffic(self, rule_words: np.ndarray) -> np.ndarray:
        """Per rule: [support bytes, pair bytes, support nodes, pair nodes]."""
        w = np.ascontiguousarray(rule_words, np.int32)
        out = np.zeros((w.shape[0], 4), np.int64)
        if self.lib.gpu_plan_traffic(self.ctx, w.ctypes.data, int(w.shape[0]), out.ctypes.data) != 0:
            raise RuntimeError("gpu_plan_traffic failed")
        return out

    def set_tuning(self, key: str, value: int):
        if self.lib.gpu_set_tuning(self.ctx, key.encode(), int(value)) != 0:
            raise KeyError(key)

    def get_tuning(self, key: str) -> int:
        return int(self.lib.gpu_get_tuning(self.ctx, key.encode()))

    def close(self):
        if self.ctx:
            self.lib.gpu_free_graph(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ResidentBatch:
    """gpu_batch_create / gpu_batch_run / gpu_batch_free: rules compiled once,
    kept in HBM, evaluated by replaying the batch's HIP graph."""

    def __init__(self, engine: Engine, rule_words: np.ndarray, flags: int = 0, min_support=None,
                 min_head_coverage: float = 0.0):
        self.engine = engine
        w = np.ascontiguousarray(rule_words, np.int32)
        self.n = int(w.shape[0])
        if min_support is None and not min_head_coverage:
            self.h = engine.lib.gpu_batch_create(engine.ctx, w.ctypes.data, self.n, int(flags))
        else:   # gpu_batch_create_min: support thresholds gate the body-pair phase
            self.h = engine.lib.gpu_batch_create_min(engine.ctx, w.ctypes.data, self.n, int(flags),
                                                     int(1 if min_support is None else min_support),
                                                     float(min_head_coverage))
        if not self.h:
            raise RuntimeError("gpu_batch_create failed")
        self.out = np.zeros(self.n, RESULT64_DTYPE)

    def run(self) -> np.ndarray:
        if self.engine.lib.gpu_batch_run(self.h, self.out.ctypes.data) != 0:
            raise RuntimeError("gpu_batch_run failed")
        return self.out

    def close(self):
        if self.h:
            self.engine.lib.gpu_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rule_words_for_graph(rules, g):
    """Structured rules over a workloads.Graph loaded with Engine.from_graph."""
    from . import workloads as W
    pred_id = np.arange(g.n_predicates, dtype=np.int64) + g.n_entities
    return W.encode_rules(rules, None, pred_id)

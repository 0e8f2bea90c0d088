"""CPU oracle for the counting path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package; the product library never does. `Oracle` wraps oracle.cpp
(compiled by oracle/Makefile into oracle/_build/liboracle.so) through ctypes.
Each result row is (support, headSize, headSupport, bodyPairs, bodySize,
skipped, nodes); see oracle.cpp for the semantics and their reference
citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

FIELDS = ("support", "headSize", "headSupport", "bodyPairs", "bodySize", "skipped", "nodes")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "-f", "Makefile"])
    return LIB_PATH


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_graph_new.restype = ctypes.c_void_p
        lib.oracle_graph_new.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64]
        lib.oracle_graph_free.argtypes = [ctypes.c_void_p]
        lib.oracle_eval.restype = ctypes.c_int
        lib.oracle_eval.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        lib.oracle_eval_support.restype = ctypes.c_int
        lib.oracle_eval_support.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        _lib = lib
    return _lib


class Oracle:
    def __init__(self, s, p, o):
        lib = _load()
        self._s = np.ascontiguousarray(s, np.int32)
        self._p = np.ascontiguousarray(p, np.int32)
        self._o = np.ascontiguousarray(o, np.int32)
        self._h = lib.oracle_graph_new(self._s.ctypes.data, self._p.ctypes.data, self._o.ctypes.data,
                                       int(self._s.shape[0]))

    def eval(self, rule_words: np.ndarray) -> np.ndarray:
        w = np.ascontiguousarray(rule_words, np.int32)
        out = np.zeros((w.shape[0], 7), np.int64)
        _load().oracle_eval(self._h, w.ctypes.data, int(w.shape[0]), out.ctypes.data)
        return out

    def eval_support(self, rule_words: np.ndarray) -> np.ndarray:
        """eval() without the body-pair search (bodyPairs = bodySize = 0)."""
        w = np.ascontiguousarray(rule_words, np.int32)
        out = np.zeros((w.shape[0], 7), np.int64)
        _load().oracle_eval_support(self._h, w.ctypes.data, int(w.shape[0]), out.ctypes.data)
        return out

    def close(self):
        if self._h:
            _load().oracle_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Serving driver: continuous batching through the library's server
(gpu_serve_* in include/gpu_rules_api.h) and a Poisson load generator.

Latency is measured from each request's scheduled arrival time to the
moment its result is available (`completeNs`), both on the library's
monotonic clock (gpu_monotonic_ns, the same CLOCK_MONOTONIC that
time.monotonic_ns reads on Linux).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import lib as L

COMPLETION_DTYPE = np.dtype([("tag", "<u8"), ("result", L.RESULT64_DTYPE), ("submitNs", "<i8"),
                             ("admitNs", "<i8"), ("completeNs", "<i8"), ("batchSize", "<i4"), ("pad", "<i4")])

_SIGS = {
    "gpu_serve_start": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                          ctypes.c_int32]),
    "gpu_serve_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "gpu_serve_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "gpu_serve_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_serve_stop": (None, [ctypes.c_void_p]),
    "gpu_serve_slot_bytes": (ctypes.c_int64, [ctypes.c_void_p]),
    "gpu_serve_heavy_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_monotonic_ns": (ctypes.c_int64, []),
}
L._SIGS.update(_SIGS)

STAT_NAMES = ("submitted", "admitted", "completed", "batches", "refused", "deferred_for_memory", "largest_batch",
              "peak_in_flight")


def _lib():
    lib = L.load_library()
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class Server:
    def __init__(self, engine: L.Engine, max_batch: int = 4096, max_in_flight: int = 3, mem_budget: int = 0,
                 flags: int = L.UNGATED):
        self.lib = _lib()
        self.engine = engine
        self.h = self.lib.gpu_serve_start(engine.ctx, int(max_batch), int(max_in_flight), int(mem_budget), int(flags))
        if not self.h:
            raise RuntimeError("gpu_serve_start failed")
        self._buf = np.zeros(65536, COMPLETION_DTYPE)

    def submit(self, rule_words: np.ndarray, tags: np.ndarray) -> int:
        w = np.ascontiguousarray(rule_words, np.int32)
        t = np.ascontiguousarray(tags, np.uint64)
        n = int(w.shape[0])
        if n == 0:
            return 0
        rc = self.lib.gpu_serve_submit(self.h, w.ctypes.data, t.ctypes.data, n)
        if rc != n:
            raise RuntimeError("gpu_serve_submit returned %d" % rc)
        return rc

    def poll(self, wait_ms: int = 0) -> np.ndarray:
        k = self.lib.gpu_serve_poll(self.h, self._buf.ctypes.data, int(self._buf.shape[0]), int(wait_ms))
        if k < 0:
            raise RuntimeError("gpu_serve_poll failed")
        return self._buf[:k].copy()

    def stats(self) -> dict:
        s = np.zeros(8, np.int64)
        rc = self.lib.gpu_serve_stats(self.h, s.ctypes.data)
        d = dict(zip(STAT_NAMES, (int(x) for x in s)))
        d["device_failure"] = rc == 1
        h = np.zeros(2, np.int64)
        if self.lib.gpu_serve_heavy_stats(self.h, h[0:].ctypes.data, h[1:].ctypes.data) == 0:
            d["heavy_requests"], d["heavy_batches"] = int(h[0]), int(h[1])
        return d

    def slot_bytes(self) -> int:
        """HBM each execution slot holds for the server's life."""
        return int(self.lib.gpu_serve_slot_bytes(self.h))

    def stop(self):
        if self.h:
            self.lib.gpu_serve_stop(self.h)
            self.h = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def now_ns() -> int:
    return int(_lib().gpu_monotonic_ns())


def run_stream(server: Server, rule_words: np.ndarray, arrivals_s: np.ndarray, lead_s: float = 0.005,
               progress=None, progress_every_s: float = 10.0) -> dict:
    """Submits request i at time t0 + arrivals_s[i] (pacing on the host),
    collects every completion, and returns per-request latencies (ms, from
    the scheduled arrival) with throughput. `progress(submitted, completed)`
    is called about every `progress_every_s` seconds."""
    n = int(rule_words.shape[0])
    last_report = time.time()
    t0 = now_ns() + int(lead_s * 1e9)
    sched = t0 + (np.asarray(arrivals_s, np.float64) * 1e9).astype(np.int64)
    done = np.zeros(n, COMPLETION_DTYPE)
    got = np.zeros(n, bool)
    i = 0
    ndone = 0
    tags = np.arange(n, dtype=np.uint64)
    while ndone < n:
        if progress is not None and time.time() - last_report >= progress_every_s:
            progress(i, ndone)
            last_report = time.time()
        now = now_ns()
        j = int(np.searchsorted(sched, now, side="right"))
        if j > i:
            server.submit(rule_words[i:j], tags[i:j])
            i = j
        c = server.poll(0)
        if c.shape[0]:
            idx = c["tag"].astype(np.int64)
            done[idx] = c
            got[idx] = True
            ndone += c.shape[0]
            continue
        if i < n:
            wait = (sched[i] - now_ns()) / 1e9
            if wait > 0.0002:
                time.sleep(min(wait, 0.002))
        else:
            c = server.poll(2)
            if c.shape[0]:
                idx = c["tag"].astype(np.int64)
                done[idx] = c
                got[idx] = True
                ndone += c.shape[0]
    lat_ms = (done["completeNs"] - sched) / 1e6
    span_s = (done["completeNs"].max() - sched.min()) / 1e9
    return {"completions": done, "latency_ms": lat_ms, "span_s": span_s, "rules_per_s": n / span_s}

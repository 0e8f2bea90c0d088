"""Serving loop on the GPU: every result is the same alone, mixed into
batches, and through the plain batch API (exact counts: tolerance 0); HBM
admission defers and refuses requests without changing results; a Poisson
stream completes with every request accounted for."""
import sys

import numpy as np
import pytest

from masterthesis_gpu_acceleration_amd import lib as L
from masterthesis_gpu_acceleration_amd import serving as S

import goldens

pytestmark = pytest.mark.gpu

FIELDS = ("support", "headSize", "headSupport", "bodySize", "bodyPairs", "status")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    fx, g, rules = goldens.load("dbpedia_shaped_5k")
    eng = L.Engine.from_graph(g)
    words = L.rule_words_for_graph(rules, g)
    ref = eng.evaluate(words, L.UNGATED)
    yield eng, words, ref
    eng.close()


def same(a, b):
    return all(np.array_equal(a[f], b[f]) for f in FIELDS)


def collect(server, n, timeout_s=60):
    out = np.zeros(n, S.COMPLETION_DTYPE)
    got = 0
    import time
    t0 = time.time()
    while got < n:
        c = server.poll(50)
        out[c["tag"].astype(np.int64)] = c
        got += c.shape[0]
        assert time.time() - t0 < timeout_s, "serving did not complete"
    return out


def test_alone_and_mixed_match_batch(setup):
    say("serve alone/mixed")
    eng, words, ref = setup
    srv = S.Server(eng, max_batch=256, max_in_flight=3)
    try:
        n = 60
        alone = np.zeros(n, S.COMPLETION_DTYPE)
        for i in range(n):                     # one request at a time
            srv.submit(words[i:i + 1], np.array([i], np.uint64))
            c = srv.poll(5000)
            while c.shape[0] == 0:
                c = srv.poll(5000)
            alone[i] = c[0]
        assert np.all(alone["batchSize"] == 1)
        srv.submit(words, np.arange(words.shape[0], dtype=np.uint64))   # all at once
        mixed = collect(srv, words.shape[0])
        st = srv.stats()
        assert st["batches"] >= 2 and st["completed"] == n + words.shape[0] and not st["device_failure"]
    finally:
        srv.stop()
    assert same(alone["result"], ref[:n])
    assert same(mixed["result"], ref)
    assert (mixed["batchSize"] > 1).any()


def test_memory_admission(setup):
    say("serve admission")
    eng, words, ref = setup
    # budget: what the two slots hold for their life plus room for about one small request
    probe = S.Server(eng, max_batch=64, max_in_flight=2)
    slot = probe.slot_bytes()
    probe.stop()
    assert slot > 0
    srv = S.Server(eng, max_batch=64, max_in_flight=2, mem_budget=2 * slot + 200_000)
    try:
        srv.submit(words[:80], np.arange(80, dtype=np.uint64))
        out = collect(srv, 80)
        st = srv.stats()
    finally:
        srv.stop()
    ok = out["result"]["status"] == 0
    refused = out["result"]["status"] == 3
    assert (ok | refused).all()
    assert st["deferred_for_memory"] > 0 or st["refused"] > 0
    assert same(out["result"][ok], ref[:80][ok])


def test_poisson_stream(setup):
    say("serve poisson")
    eng, words, ref = setup
    n = 2000
    idx = np.arange(n) % words.shape[0]
    rng = np.random.default_rng(3)
    arrivals = np.cumsum(rng.exponential(1.0 / 20000.0, n))
    srv = S.Server(eng, max_batch=1024, max_in_flight=3)
    try:
        r = S.run_stream(srv, words[idx], arrivals)
        st = srv.stats()
    finally:
        srv.stop()
    assert st["completed"] == n and not st["device_failure"]
    assert same(r["completions"]["result"], ref[idx])
    assert (r["latency_ms"] > 0).all()
    say("poisson: %.0f rules/s, p50 %.3f ms, p99 %.3f ms, batches %d" % (
        r["rules_per_s"], np.percentile(r["latency_ms"], 50), np.percentile(r["latency_ms"], 99), st["batches"]))


def test_heavy_lane(setup):
    """Requests whose estimated work passes serve_heavy_work run in heavy
    batches of their own on a reduced grid; every result is unchanged."""
    say("serve heavy lane")
    eng, words, ref = setup
    fx, g, rules = goldens.load("dbpedia_shaped_5k")
    est = L.HostGraph(g).work(words)
    assert est.shape[0] == words.shape[0] and (est >= 0).all()
    thr = int(np.percentile(est[est > 0], 80))
    n = words.shape[0]
    try:
        eng.set_tuning("serve_heavy_work", thr)
        eng.set_tuning("serve_heavy_percent", 10)
        srv = S.Server(eng, max_batch=256, max_in_flight=3)
        try:
            srv.submit(words, np.arange(n, dtype=np.uint64))
            out = collect(srv, n)
            st = srv.stats()
        finally:
            srv.stop()
    finally:
        eng.set_tuning("serve_heavy_work", 10 ** 11)
        eng.set_tuning("serve_heavy_percent", 50)
    say("  threshold %d: %d heavy requests in %d heavy batches of %d" % (
        thr, st["heavy_requests"], st["heavy_batches"], st["batches"]))
    assert st["completed"] == n and not st["device_failure"]
    assert st["heavy_requests"] == int((est > thr).sum()) and st["heavy_batches"] >= 1
    assert same(out["result"], ref)

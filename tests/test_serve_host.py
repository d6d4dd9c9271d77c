"""Host side of the serving loop (CPU only): the request stream and the
per-request inputs.

The stream is seeded and Poisson (exponential gaps with the stated mean);
each request's inputs are built exactly as the 3-cluster workload recipe
says, falling back to the rank <= 1 space above the size caps; building the
same request twice gives identical arrays.
"""
import numpy as np

import threading

from fermicg_amd.serve import FULL_CAP, Request, _ReadyPool, build_inputs, request_stream, summarize


def test_request_stream_seeded_poisson():
    a = request_stream(2000, rate=50.0, seed=3)
    b = request_stream(2000, rate=50.0, seed=3)
    assert [(r.seed, r.n_clusters, r.M, r.R, r.arrival) for r in a] == \
           [(r.seed, r.n_clusters, r.M, r.R, r.arrival) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0          # exponential with mean 1/rate
    assert abs(np.std(gaps) / gaps.mean() - 1.0) < 0.1      # coefficient of variation 1
    assert {r.n_clusters for r in a} == set(range(2, 9))
    assert {r.M for r in a} == {4, 8, 16}
    assert {r.R for r in a} == {1, 2, 3, 4}


def test_build_inputs_recipe_and_determinism():
    small = build_inputs(Request(0, 5, 3, 4, 2))
    assert small["kind"] == "full" and small["n_var"] == 4672     # 73 Fock blocks x <= 64
    big = build_inputs(Request(1, 5, 5, 16, 1))
    assert big["kind"] == "rank1"
    st = big["state"]
    # rank <= 1: at most one cluster away from its lowest state
    assert np.all((st.cfg > 1).sum(axis=1) <= 1)
    again = build_inputs(Request(1, 5, 5, 16, 1))
    assert np.array_equal(again["state"].cfg, st.cfg)
    assert np.array_equal(again["state"].coeffs, st.coeffs)
    assert np.array_equal(again["fops"].gdata, big["fops"].gdata)
    assert FULL_CAP == 2_000_000


def test_summarize():
    res = [{"latency": 1.0, "done": 2.0, "arrival": 1.0, "error": None, "converged": True},
           {"latency": 3.0, "done": 5.0, "arrival": 2.0, "error": None, "converged": True}]
    s = summarize(res)
    assert s["requests"] == 2 and s["converged"] == 2
    assert abs(s["span_s"] - 4.0) < 1e-12 and abs(s["solves_per_s"] - 0.5) < 1e-12
    assert s["p50_latency_s"] == 2.0


def test_ready_pool_order():
    # smallest n_var x R first (ties by arrival), slot 0 the largest; fifo by
    # readiness; one None per slot ends the slots after the requests
    def item(rid, n_var, R, arrival):
        return (Request(rid, 0, 2, 4, R, arrival), {"n_var": n_var}, 0.0)
    items = [item(0, 100, 2, 0.0), item(1, 50, 1, 0.1), item(2, 50, 1, 0.05),
             item(3, 1000, 4, 0.2), item(4, 10, 1, 0.3)]
    pool = _ReadyPool("size")
    for it in items:
        pool.put(it)
    pool.put(None)
    pool.put(None)
    got = [pool.get(largest=True)[0].rid, pool.get()[0].rid, pool.get()[0].rid,
           pool.get()[0].rid, pool.get(largest=True)[0].rid]
    assert got == [3, 4, 2, 1, 0]
    assert pool.get() is None and pool.get(largest=True) is None
    fifo = _ReadyPool("fifo")
    for it in items:
        fifo.put(it)
    assert [fifo.get(largest=k % 2 == 0)[0].rid for k in range(5)] == [0, 1, 2, 3, 4]
    # a waiting slot wakes on put
    out = []
    t = threading.Thread(target=lambda: out.append(fifo.get()))
    t.start()
    fifo.put(items[4])
    t.join(5)
    assert out and out[0][0].rid == 4

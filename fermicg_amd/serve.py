"""Serving loop: a stream of independent Davidson solves on one GPU.

Requests arrive as a Poisson stream.  Each names a seeded synthetic system:
4*n_clusters orbitals at half filling in n_clusters clusters of 4, M states
per Fock sector, R roots.  Its integrals, cluster bases and variational
space are built as for the 3-cluster TPSCI workload: every configuration of
the Fock blocks within 2 transferred electrons when that is at most 2e6
configurations and the dense bound on the Hamiltonian blocks between
connected Fock blocks is at most 2e9 elements, else every rank <= 1
configuration of those blocks.

Continuous batching:
  * producer processes (spawned, CPU only) build each request's inputs as it
    arrives: local eigenbases, ClusterOps, ClusteredTerms, space;
  * `slots` solver threads, each with its own library context (HIP stream),
    take ready requests, so solves of different sizes run concurrently on
    the device; the library releases the GIL in every call.  A free slot
    takes the request that became ready first (order="fifo"), or with
    order="size" slot 0 takes the largest ready request (n_var x R) and the
    others the smallest (measured no better on the Poisson mix);
  * admission by HBM footprint: the library holds a process-wide budget for
    live plans (fcg_set_plan_budget); a plan reserves its device bytes
    before allocating anything and is refused when they do not fit beside
    the resident plans, in which case the request waits until a running
    solve releases its plan;
  * plan builds and solves of different requests overlap.

Each result records arrival, ready (inputs built), start, plan and solve
times; latency = completion - arrival.  Solves are deterministic (fixed-order
reductions everywhere), so a request's energies are the same alone and in
the mix.
"""
import concurrent.futures as cf
import multiprocessing as mp
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _lib

FULL_CAP = 2_000_000          # configurations
FULL_MAX_ELEMS = 2_000_000_000  # dense bound on Hamiltonian-block elements (16 GB in FP64)


@dataclass
class Request:
    rid: int
    seed: int
    n_clusters: int
    M: int
    R: int
    arrival: float = 0.0


def request_stream(n, rate=50.0, seed=0):
    """n requests with exponential inter-arrival times (mean 1/rate s),
    n_clusters uniform in 2..8, M uniform in {4, 8, 16}, R uniform in 1..4."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        out.append(Request(i, int(rng.integers(1, 2 ** 31 - 1)), int(rng.integers(2, 9)),
                           int(rng.choice([4, 8, 16])), int(rng.integers(1, 5)), float(t[i])))
    return out


def build_inputs(req):
    """Host inputs of one request (runs in a producer process)."""
    from .marshal import flatten_ham, flatten_ops
    from .state import TPSCIstate
    from .workloads import build_problem, rank1_space
    N = req.n_clusters
    t0 = time.perf_counter()
    p = build_problem(4 * N, N, req.M, 8 * N, req.seed, max_transfer=2)
    dims = {f: int(np.prod([p.bases[c].n_states(f[c]) for c in range(N)])) for f in p.focks}
    full = sum(dims.values())
    if full <= FULL_CAP:
        # dense bound on the blocks a full-block plan would hold
        from .state import fock_sub
        tcs = [tc for tc, _ in p.clustered_ham]
        elems = 0
        for f in p.focks:
            for tc in tcs:
                fk = fock_sub(f, tc)
                if fk in dims:
                    elems += dims[f] * dims[fk]
        if elems > FULL_MAX_ELEMS:
            full = FULL_CAP + 1
    if full <= FULL_CAP:
        st = TPSCIstate.full_blocks(p.bases, p.focks, R=req.R)
        kind = "full"
    else:
        st = rank1_space(p, R=req.R, frac=1.0, seed=req.seed)
        kind = "rank1"
    rng = np.random.Generator(np.random.PCG64(req.seed + 1))
    v = rng.normal(size=(len(st), req.R))
    v /= np.linalg.norm(v, axis=0)
    st.set_vector(v)
    fops = flatten_ops(p.cluster_ops, [len(c) for c in p.clusters])
    fham = flatten_ham(p.clustered_ham)
    return {"fops": fops, "fham": fham, "state": st, "kind": kind, "n_var": len(st),
            "n_fock": len(p.focks), "build_s": time.perf_counter() - t0}


class SolveServer:
    def __init__(self, slots=4, precision="fp64", producers=8, hbm_budget=200e9,
                 conv_thresh=1e-7, max_iter=300, refresh_k=0, device=0, order="fifo"):
        if order not in ("size", "fifo"):
            raise ValueError("order must be 'size' or 'fifo'")
        self.order = order
        self.slots = slots
        self.precision = precision
        self.producers = producers
        self.budget = float(hbm_budget)
        self.conv_thresh = conv_thresh
        self.max_iter = max_iter
        self.refresh_k = refresh_k
        self.device = device
        self.precs = ("fp64",) if precision == "fp64" else ("fp64", precision)

    # -- one solve on a given context
    def _solve(self, ctx, inp, admit):
        N = inp["state"].N
        ops = _lib.DeviceOps.upload(ctx, inp["fops"])
        ham = _lib.DeviceHam.upload(ctx, inp["fham"], N)
        sp = _lib.DeviceSpace.tpsci(ctx, inp["state"])
        pl, foot = admit(ctx, ham, ops, sp)
        t1 = time.perf_counter()
        st = inp["state"]
        e, _, stats = pl.davidson(st.coeffs, conv_thresh=self.conv_thresh, max_iter=self.max_iter,
                                  max_ss_vecs=12, lindep_thresh=1e-12, precision=self.precision,
                                  refresh_k=self.refresh_k)
        t2 = time.perf_counter()
        del pl
        return e, stats, foot, t1, t2

    def solve_alone(self, req):
        """The request on its own (fresh context, no other work)."""
        inp = build_inputs(req)
        ctx = _lib.Context(self.device)

        def admit(ctx_, ham, ops, sp):
            pl, foot = _lib.Plan.create_limited(ctx_, ham, ops, sp, self.budget,
                                                precisions=self.precs)
            if pl is None:
                raise MemoryError("request %d needs %d device bytes" % (req.rid, foot))
            return pl, foot

        e, stats, foot, _, _ = self._solve(ctx, inp, admit)
        ctx.synchronize()
        return e, stats

    def run(self, requests, log=None):
        """Serve `requests` (arrival times in seconds from now); returns the
        per-request results in request order."""
        ready = _ReadyPool(self.order)
        results = {}
        mem = threading.Condition()
        state = {"in_use": 0.0, "peak": 0.0, "running": 0}

        held = threading.local()

        def admit(ctx, ham, ops, sp):
            # the library reserves the plan's footprint against the process-wide
            # budget before allocating; a plan that does not fit waits for a
            # running solve to release its plan
            while True:
                pl, foot = _lib.Plan.create_limited(ctx, ham, ops, sp, 2 ** 62,
                                                    precisions=self.precs)
                if pl is not None:
                    with mem:
                        state["in_use"] += foot
                        state["peak"] = max(state["peak"], state["in_use"])
                    held.foot = foot
                    return pl, foot
                if foot > self.budget:
                    raise MemoryError("request needs %d device bytes, budget %d"
                                      % (foot, self.budget))
                with mem:
                    mem.wait(timeout=0.5)

        def slot(k):
            ctx = _lib.Context(self.device)
            while True:
                item = ready.get(largest=(k == 0))
                if item is None:
                    break
                req, inp, t_ready = item
                with mem:
                    state["running"] += 1
                held.foot = 0
                t0 = time.perf_counter()
                try:
                    e, stats, foot, t1, t2 = self._solve(ctx, inp, admit)
                    err = None
                except Exception as ex:  # reported per request
                    e, stats, foot, t1, t2, err = None, None, held.foot, t0, time.perf_counter(), repr(ex)
                with mem:
                    state["in_use"] -= held.foot
                    state["running"] -= 1
                    mem.notify_all()
                t_done = time.perf_counter()
                results[req.rid] = {
                    "rid": req.rid, "n_clusters": req.n_clusters, "M": req.M, "R": req.R,
                    "kind": inp["kind"], "n_var": inp["n_var"],
                    "energies": None if e is None else [float(x) for x in e],
                    "converged": None if stats is None else stats["converged"],
                    "iterations": None if stats is None else stats["iterations"],
                    "arrival": req.arrival, "ready": t_ready - t_start, "start": t0 - t_start,
                    "plan_s": t1 - t0, "solve_s": t2 - t1, "done": t_done - t_start,
                    "latency": t_done - t_start - req.arrival, "footprint": foot, "error": err}
                if log:
                    log("done", req.rid, "N=%d M=%d R=%d %s n=%d lat %.2fs plan %.2fs solve %.2fs"
                        % (req.n_clusters, req.M, req.R, inp["kind"], inp["n_var"],
                           results[req.rid]["latency"], t1 - t0, t2 - t1))
            ctx.close()

        pool = self._producer_pool()
        prev_budget, _ = _lib.plan_budget()
        _lib.set_plan_budget(self.budget)
        threads = [threading.Thread(target=slot, args=(k,), daemon=True) for k in range(self.slots)]
        t_start = time.perf_counter()
        for th in threads:
            th.start()
        futs = []
        for req in sorted(requests, key=lambda r: r.arrival):
            dt = req.arrival - (time.perf_counter() - t_start)
            if dt > 0:
                time.sleep(dt)
            f = pool.submit(build_inputs, req)
            f.add_done_callback(lambda f, req=req: ready.put((req, f.result(), time.perf_counter())))
            futs.append(f)
        cf.wait(futs)
        for _ in threads:
            ready.put(None)
        for th in threads:
            th.join()
        _lib.set_plan_budget(prev_budget)
        self.peak_bytes = state["peak"]
        self.in_use_end = state["in_use"]
        return [results[r.rid] for r in requests]

    def set_precision(self, precision):
        self.precision = precision
        self.precs = ("fp64",) if precision == "fp64" else ("fp64", precision)

    def _producer_pool(self):
        """Input builders: spawned CPU-only processes, started (and warmed)
        before this process touches the GPU when used first; producers=0
        builds in threads instead."""
        if getattr(self, "_pool", None) is None:
            if self.producers > 0:
                self._pool = cf.ProcessPoolExecutor(self.producers,
                                                    mp_context=mp.get_context("spawn"))
                list(self._pool.map(_noop, range(self.producers)))
            else:
                self._pool = cf.ThreadPoolExecutor(2)
        return self._pool

    def start(self):
        """Start the input builders (call before any GPU work in the process)."""
        self._producer_pool()
        return self

    def close(self):
        if getattr(self, "_pool", None) is not None:
            self._pool.shutdown()
            self._pool = None


def _noop(i):
    return i


class _ReadyPool:
    """Requests whose inputs are built.  get() blocks until one is there and
    returns the smallest (or, with largest=True, the largest) by n_var x R,
    ties by arrival; with order="fifo" always the first that became ready.
    None (one per slot, put after every request) ends a slot."""

    def __init__(self, order):
        self.order = order
        self.items = []
        self.stops = 0
        self.cv = threading.Condition()
        self.seq = 0

    def put(self, item):
        with self.cv:
            if item is None:
                self.stops += 1
            else:
                req, inp, _ = item
                self.items.append((inp["n_var"] * req.R, req.arrival, self.seq, item))
                self.seq += 1
            self.cv.notify_all()

    def get(self, largest=False):
        with self.cv:
            while not self.items and not self.stops:
                self.cv.wait()
            if not self.items:
                self.stops -= 1
                return None
            if self.order == "fifo":
                i = min(range(len(self.items)), key=lambda j: self.items[j][2])
            elif largest:
                i = max(range(len(self.items)), key=lambda j: (self.items[j][0], -self.items[j][1]))
            else:
                i = min(range(len(self.items)), key=lambda j: self.items[j][:2])
            return self.items.pop(i)[3]


def summarize(results):
    lat = np.array([r["latency"] for r in results if r["error"] is None])
    span = max(r["done"] for r in results) - min(r["arrival"] for r in results)
    ok = sum(1 for r in results if r["error"] is None and r["converged"])
    return {"requests": len(results), "converged": ok, "solves_per_s": ok / span,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "span_s": span}

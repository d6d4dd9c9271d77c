"""One plan shared by several library contexts (streams) on one GPU.

Each context keeps its own interleaved copy of c, so sigma calls from two
contexts on one plan may overlap; both results must equal the oracle's.  A
context destroyed before the plan stops being one of its users, so
destroying the plan afterwards touches no freed context.
"""
import ctypes
import threading

import numpy as np
import pytest

from fermicg_amd import _lib
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.tpsci import Engine, cluster_sizes_of
from fermicg_amd.workloads import build_problem, rank1_space, tpsci_small
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _sigma_on(ctx, pl, c):
    R = c.shape[1]
    cc = np.ascontiguousarray(c.T)
    out = np.zeros((R, c.shape[0]))
    P = ctypes.POINTER(ctypes.c_double)
    ctx.check(ctx.L.fcg_sigma(ctx.h, pl.h, R, cc.ctypes.data_as(P), out.ctypes.data_as(P), 0),
              "fcg_sigma")
    return out.T.copy()


@pytest.mark.parametrize("case", ["full", "rank1"])
def test_two_contexts_overlap_on_one_plan(case):
    if case == "full":
        p, st = tpsci_small(seed=0, R=2)
    else:
        p = build_problem(20, 5, 8, 40, seed=3, max_transfer=2)
        st = rank1_space(p, R=2, frac=0.5, seed=3)
    rng = np.random.Generator(np.random.PCG64(9))
    cs = [rng.normal(size=(len(st), 2)) for _ in range(2)]
    orc = Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                 flatten_ham(p.clustered_ham), st)
    refs = [orc.sigma(c) for c in cs]
    eng = Engine(0)
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham)
    ctxs = [eng.ctx, _lib.Context(0)]
    errs = [0.0, 0.0]

    def worker(k):
        for _ in range(20):
            s = _sigma_on(ctxs[k], pl, cs[k])
            errs[k] = max(errs[k], np.abs(s - refs[k]).max() / np.abs(refs[k]).max())

    th = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    print(case, len(st), errs, flush=True)
    assert max(errs) < 1e-12
    ctxs[1].close()          # a user context goes first
    eng.drop_plans()         # then the plan: must not touch the freed context
    s = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham).sigma(cs[0])
    assert np.abs(s - refs[0]).max() / np.abs(refs[0]).max() < 1e-12

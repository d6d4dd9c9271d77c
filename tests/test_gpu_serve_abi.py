"""The serving loop behind the C ABI (fcg_serve_start/submit/poll/stop).

Six small requests (2-4 clusters, M 4 to 16, 1-2 roots, full-block and
rank-1 spaces) are uploaded once and submitted together to a server with
three solver slots and a 200 MB plan budget, so solves overlap on the device
and some wait for memory.  Checks:
  * every request converged, and its energies equal those of the same solve
    run alone through fcg_plan_create + fcg_davidson on a fresh context to
    1e-12 Eh (the solves are deterministic);
  * against the dense Hamiltonian of the oracle (build_full_H restatement,
    pinned to determinant FCI elsewhere) for the spaces small enough to
    diagonalise: eigenvalues within 1e-8 Eh;
  * the plans resident at any time never exceed the budget, and some solves
    did run concurrently;
  * a request whose plan is larger than the whole budget fails with
    FCG_ENOMEM instead of waiting.
"""
import numpy as np
import pytest

from fermicg_amd import _lib
from fermicg_amd.serve import Request, build_inputs

pytestmark = pytest.mark.gpu

REQS = [Request(0, 11, 3, 4, 1, 0.0), Request(1, 12, 2, 8, 2, 0.0),
        Request(2, 13, 4, 8, 2, 0.0), Request(3, 14, 3, 16, 1, 0.0),
        Request(4, 15, 3, 4, 2, 0.0), Request(5, 16, 2, 4, 2, 0.0)]
BUDGET = 2e8
OPTS = dict(conv_thresh=1e-8, max_iter=300, max_ss_vecs=12, lindep_thresh=1e-12)


@pytest.fixture(scope="module")
def uploaded():
    ctx = _lib.Context(0)
    items = []
    for req in REQS:
        inp = build_inputs(req)
        st = inp["state"]
        items.append((req, inp, _lib.DeviceHam.upload(ctx, inp["fham"], st.N),
                      _lib.DeviceOps.upload(ctx, inp["fops"]), _lib.DeviceSpace.tpsci(ctx, st)))
    ctx.synchronize()
    yield ctx, items
    del items
    ctx.close()


def test_served_equals_alone_and_dense(uploaded):
    ctx, items = uploaded
    srv = _lib.Server(slots=3, hbm_budget=BUDGET, **OPTS)
    tickets = [srv.submit(ham, ops, sp, inp["state"].coeffs) for _, inp, ham, ops, sp in items]
    res = [srv.poll(t) for t in tickets]
    srv.stop()
    assert _lib.plan_budget()[1] == 0
    for (req, inp, ham, ops, sp), r in zip(items, res):
        print(req.rid, inp["kind"], inp["n_var"], r["rc"], r["energies"], r["stats"]["iterations"],
              r["times"], flush=True)
        assert r["rc"] == 0 and r["stats"]["converged"]
        alone_ctx = _lib.Context(0)
        pl = _lib.Plan.create(alone_ctx, ham, ops, sp)
        e, _, st = pl.davidson(inp["state"].coeffs, **OPTS)
        del pl
        alone_ctx.close()
        assert np.abs(r["energies"] - e).max() < 1e-12
        if inp["n_var"] <= 3000:
            from oracle.pyoracle import Oracle
            w = np.linalg.eigvalsh(Oracle(inp["fops"], inp["fham"], inp["state"]).full_H())
            assert np.abs(r["energies"] - w[:req.R]).max() < 1e-8
    # plans resident at each admission within the budget; some solves overlapped
    iv = [(r["times"]["planned"], r["times"]["finished"], r["times"]["plan_bytes"]) for r in res]
    for t, _, _ in iv:
        held = sum(b for a, f, b in iv if a <= t < f)
        assert held <= BUDGET
    overlap = any(a1 < f2 and a2 < f1 for i, (a1, f1, _) in enumerate(iv)
                  for (a2, f2, _) in iv[i + 1:])
    assert overlap


def test_oversized_request_fails(uploaded):
    ctx, items = uploaded
    req, inp, ham, ops, sp = items[3]
    srv = _lib.Server(slots=1, hbm_budget=1024, **OPTS)
    t = srv.submit(ham, ops, sp, inp["state"].coeffs)
    r = srv.poll(t)
    srv.stop()
    print(r["rc"], r["times"], flush=True)
    assert r["rc"] == -2 and r["times"]["plan_bytes"] > 1024

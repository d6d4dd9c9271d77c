"""Serving loop: concurrent solves give the same results as solves alone.

Six small requests (2-4 clusters, M 4 to 16, 1-2 roots, full-block and
rank-1 spaces, plans of 1e5 to 1.4e8 device bytes) run through SolveServer
with three solver slots and a 200 MB HBM budget, so solves overlap on the
device and some wait for memory; each
request is then solved alone on a fresh context.  Bar: energies equal to
1e-12 Eh (the solves are deterministic), every solve converged, and the
admission accounting returns to zero.
"""
import numpy as np
import pytest

from fermicg_amd.serve import Request, SolveServer, summarize

pytestmark = pytest.mark.gpu

REQS = [Request(0, 11, 3, 4, 1, 0.00), Request(1, 12, 2, 8, 2, 0.01),
        Request(2, 13, 4, 8, 2, 0.02), Request(3, 14, 3, 16, 1, 0.02),
        Request(4, 15, 3, 4, 2, 0.03), Request(5, 16, 2, 4, 2, 0.05)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mix_equals_alone(precision):
    srv = SolveServer(slots=3, precision=precision, producers=0, hbm_budget=2e8,
                      conv_thresh=1e-8)
    res = srv.run(REQS, log=lambda *a: print(*a, flush=True))
    print(summarize(res), "peak bytes", srv.peak_bytes, flush=True)
    assert all(r["error"] is None and r["converged"] for r in res)
    assert srv.in_use_end == 0 and 0 < srv.peak_bytes <= 2e8
    for req, r in zip(REQS, res):
        e_alone, st = srv.solve_alone(req)
        print(req.rid, r["kind"], r["n_var"], r["energies"], list(e_alone), flush=True)
        assert np.abs(np.array(r["energies"]) - e_alone).max() < 1e-12
    srv.close()

"""The benchmark configurations at their full sizes on the GPU.

* 24 orbitals / 6 clusters / M=16 / Tucker ranks 1-8 (bst_medium, 464,878
  variables, 2 roots): BST sigma in FP64, FP32 and split-BF16 against the C
  restatement of build_sigma! (oracle/fcg_oracle_bst.c) on 2,000 seeded bra
  Tucker blocks; Davidson FP64 against mixed FP32 with FP64 refresh.
* 32 orbitals / 8 clusters / M=16 / rank-1 (tpsci_sparse, 2,000,089
  configurations, 4 roots): TPSCI sigma in FP64, FP32 and split-BF16 against
  the C restatement of tps_ci_matvec on 2,000 seeded rows; Davidson FP64
  against mixed FP32.
* A served request at the top of the serving mix (8 clusters, M=16, 2 roots,
  rank-1 space of 580,673 configurations): its sigma against the oracle on
  sampled rows, and its energies in a concurrent mix equal to its energies
  alone.
* 48 orbitals / 12 clusters / M=32 / rank-1 (pt2_12c_M32, 9,020,329
  configurations): the sparse PT2 at thresh_foi 1e-5 against the oracle's
  restatement over the configurations reached, with the first-order state
  cut to 12 nonzero coefficients so that the oracle finishes in seconds.

Bars: sigma relative to max|sigma| 1e-12 (FP64), 1e-5 (FP32), 1e-4
(split-BF16); mixed-precision energies within 1 uEh of FP64 (converged on
FP64 residuals of 1e-8); mix equals alone to 1e-12 Eh.
"""
import os

import numpy as np
import pytest

from fermicg_amd import _lib
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.tpsci import Engine, cluster_sizes_of
from oracle.pyoracle import BstOracle, Oracle

pytestmark = pytest.mark.gpu

PRECS = ("fp64", "fp32", "bf16x2", "fp32_bf16c")
# fp32_bf16c: c rounded to bf16 (unit roundoff 2^-9) in the one-row apply
TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x2": 1e-4, "fp32_bf16c": 1e-2}
THREADS = max(1, min(16, os.cpu_count() or 1))


def _flat(p):
    return (flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
            flatten_ham(p.clustered_ham))


def _check_sigma(pl, c, idx, ref):
    scale = np.abs(ref).max()
    for prec in PRECS:
        s = pl.sigma(c, prec)
        err = np.abs(s[idx] - ref).max() / scale
        print(prec, "rel err", err, flush=True)
        assert err < TOL[prec]


def _check_mixed_solve(pl, c):
    e = {}
    for prec in ("fp64", "fp32", "fp32_bf16c"):
        ev, _, st = pl.davidson(c, conv_thresh=1e-8, max_iter=600, max_ss_vecs=12,
                                precision=prec)
        print(prec, ev, st["iterations"], st["converged"], flush=True)
        assert st["converged"]
        e[prec] = ev
    assert np.abs(e["fp32"] - e["fp64"]).max() <= 1e-6
    assert np.abs(e["fp32_bf16c"] - e["fp64"]).max() <= 1e-6


def test_bst_24o_6c_M16_full_size():
    from fermicg_amd.bst import cache_hamiltonian
    from fermicg_amd.workloads import bst_medium
    p, st = bst_medium()
    print("bst_medium", len(st), "variables", len(st.blocks), "blocks", flush=True)
    eng = Engine(0)
    pl = cache_hamiltonian(st, st, p.cluster_ops, p.clustered_ham, precisions=PRECS, engine=eng)
    fo, fh = _flat(p)
    rng = np.random.Generator(np.random.PCG64(12345))
    rows = np.sort(rng.choice(len(st.blocks), 2000, replace=False))
    ref, idx = BstOracle(fo, fh, st).sigma_rows(st.coeffs, rows, nthreads=THREADS)
    _check_sigma(pl, st.coeffs, idx, ref)
    _check_mixed_solve(pl, st.coeffs)
    del pl
    eng.clear()


def test_tpsci_32o_8c_M16_rank1_full_size():
    from fermicg_amd.workloads import tpsci_sparse
    p, st = tpsci_sparse()
    print("tpsci_sparse", len(st), "configurations", flush=True)
    assert len(st) == 2000089
    eng = Engine(0)
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham, precisions=PRECS)
    fo, fh = _flat(p)
    rng = np.random.Generator(np.random.PCG64(12345))
    idx = np.sort(rng.choice(len(st), 2000, replace=False))
    ref = Oracle(fo, fh, st).sigma_rows(st.coeffs, idx, nthreads=THREADS)
    _check_sigma(pl, st.coeffs, idx, ref)
    _check_mixed_solve(pl, st.coeffs)
    del pl
    eng.clear()


def test_served_top_request_8c_M16():
    from fermicg_amd.serve import Request, SolveServer, build_inputs
    big = Request(0, 4242, 8, 16, 2, 0.0)
    inp = build_inputs(big)
    st = inp["state"]
    print("served request", inp["kind"], inp["n_var"], flush=True)
    assert inp["kind"] == "rank1" and inp["n_var"] == 580673
    ctx = _lib.Context(0)
    ops = _lib.DeviceOps.upload(ctx, inp["fops"])
    ham = _lib.DeviceHam.upload(ctx, inp["fham"], st.N)
    sp = _lib.DeviceSpace.tpsci(ctx, st)
    pl = _lib.Plan.create(ctx, ham, ops, sp, precisions=("fp64",))
    rng = np.random.Generator(np.random.PCG64(4))
    idx = np.sort(rng.choice(len(st), 2000, replace=False))
    ref = Oracle(inp["fops"], inp["fham"], st).sigma_rows(st.coeffs, idx, nthreads=THREADS)
    s = pl.sigma(st.coeffs)
    err = np.abs(s[idx] - ref).max() / np.abs(ref).max()
    print("fp64 rel err", err, flush=True)
    assert err < 1e-12
    del pl, sp, ham, ops
    ctx.close()
    # in a concurrent mix with smaller requests, the same energies as alone
    reqs = [big, Request(1, 12, 3, 8, 2, 0.0), Request(2, 13, 4, 8, 1, 0.01)]
    srv = SolveServer(slots=3, precision="fp32", producers=0, hbm_budget=200e9,
                      conv_thresh=1e-8)
    res = srv.run(reqs, log=lambda *a: print(*a, flush=True))
    assert all(r["error"] is None and r["converged"] for r in res)
    e_alone, _ = srv.solve_alone(big)
    print("mix", res[0]["energies"], "alone", list(e_alone), flush=True)
    assert np.abs(np.array(res[0]["energies"]) - e_alone).max() < 1e-12
    srv.close()


def test_pt2_48o_12c_M32_full_size():
    """The 12-cluster PT2 workload at its full size (48 orbitals, M=32,
    9,020,329 rank <= 1 variational configurations, thresh_foi 1e-5) with the
    first-order state of the benchmark cut to its 12 largest coefficients:
    the sparse GPU PT2 against the oracle's restatement over the
    configurations reached (orc_tpsci_pt2_sparse, 16 host threads): e2 to
    1e-11 relative and the number of first-order configurations exactly;
    the mixed FP32 contraction within 1e-6 relative of FP64."""
    from fermicg_amd.tpsci import compute_pt2_energy
    from fermicg_amd.workloads import cmf_first_order_state, pt2_12c_M32
    p, st = pt2_12c_M32(M=32)
    eng = Engine(0)
    c = cmf_first_order_state(p, st, eng, select=0.1)
    eng.drop_plans()
    order = np.argsort(-np.abs(c[:, 0]), kind="stable")
    c[order[12:]] = 0.0
    st.set_vector(c / np.linalg.norm(c))
    e0 = np.array([-30000.0])
    kw = dict(E0=e0, thresh_foi=1e-5, verbose=0, method="sparse", engine=eng)
    e2 = compute_pt2_energy(st, p.cluster_ops, p.clustered_ham, **kw)
    stats = dict(compute_pt2_energy.last_stats)
    e32 = compute_pt2_energy(st, p.cluster_ops, p.clustered_ham, precision="fp32", **kw)
    orc = Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                 flatten_ham(p.clustered_ham), st)
    ref, _, nx = orc.pt2_sparse(st.coeffs, e0, thresh=1e-5, nthreads=16)
    print(e2, e32, ref, nx, stats, flush=True)
    assert stats["n_first_order"] == nx and nx > 1e7
    assert np.abs(e2 - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(e32 - e2).max() < 1e-6 * np.abs(e2).max()

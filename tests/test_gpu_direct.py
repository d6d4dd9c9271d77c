"""tps_ci_direct (src/tpsci_outer.jl:165-305) and build_full_H on the GPU.

Oracle: the C restatement of build_full_H (oracle/fcg_oracle.c, pinned to
FCI in test_fci_limit.py).  Bars: the GPU H within 1e-12 relative of the
oracle's (max-norm, and symmetric to the same bar); energies within 1e-10
Eh of the dense eigenvalues; eigenvector residuals below 1e-8 (the
Davidson path, conv_thresh 1e-9) or 1e-10 with orthonormal vectors (the
whole-matrix path, at most 500 configurations).
"""
import numpy as np
import pytest

from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.tpsci import Engine, build_full_H, cluster_sizes_of, tps_ci_direct
from fermicg_amd.workloads import tpsci_small
from oracle.pyoracle import Oracle
from tests.pt2_cases import three_clusters_m4

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


def test_direct_whole_matrix(eng):
    case = three_clusters_m4(R=2)
    st = case.state
    assert len(st) <= 500
    H_ref = Oracle(flatten_ops(case.ops, [len(c) for c in case.clusters]), flatten_ham(case.H),
                   st).full_H()
    e, v, H = tps_ci_direct(st, case.ops, case.H, engine=eng)
    print("n", len(st), "H rel", np.abs(H - H_ref).max() / np.abs(H_ref).max(), flush=True)
    assert np.abs(H - H_ref).max() < 1e-12 * np.abs(H_ref).max()
    assert np.abs(H - H.T).max() < 1e-12 * np.abs(H_ref).max()
    w = np.linalg.eigvalsh(H_ref)
    assert np.abs(e - w[:2]).max() < 1e-10
    r = H_ref @ v.coeffs - v.coeffs * e
    assert np.abs(r).max() < 1e-10
    assert np.abs(v.coeffs.T @ v.coeffs - np.eye(2)).max() < 1e-12


def test_direct_davidson_path(eng):
    p, st = tpsci_small(seed=0, R=2)
    assert len(st) > 500
    H_ref = Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                   flatten_ham(p.clustered_ham), st).full_H()
    H = build_full_H(st, p.cluster_ops, p.clustered_ham, engine=eng)
    print("n", len(st), "H rel", np.abs(H - H_ref).max() / np.abs(H_ref).max(), flush=True)
    assert np.abs(H - H_ref).max() < 1e-12 * np.abs(H_ref).max()
    e, v, H2 = tps_ci_direct(st, p.cluster_ops, p.clustered_ham, conv_thresh=1e-9, max_iter=300,
                             engine=eng)
    assert np.array_equal(H2, H)
    w = np.linalg.eigvalsh(H_ref)[:2]
    print("energies", e, w, flush=True)
    assert np.abs(e - w).max() < 1e-10
    r = H_ref @ v.coeffs - v.coeffs * e
    assert np.abs(r).max() < 1e-8

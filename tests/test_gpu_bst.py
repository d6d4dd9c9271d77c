"""GPU parity of the BST sigma build and ci_solve through the C ABI.

Oracles: U' H_TPS U (TPSCI oracle projected through the Tucker factors) on
small systems, the C restatement of build_sigma! on sampled bra blocks of a
6-cluster system.  Bars as for TPSCI: FP64 1e-12 relative, energies 1e-10
Eh against dense diagonalisation, mixed precision |dE| <= 1e-6 Eh.
"""
import numpy as np
import pytest

from fermicg_amd.bst import BSTstate, bst_full_identity, bst_pq_state, build_sigma, ci_solve, embedding
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import Engine, cluster_sizes_of, tps_ci_matvec
from fermicg_amd.workloads import build_problem, bst_medium
from oracle.pyoracle import BstOracle, Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _small(no, nc, M, mq, mr):
    p = build_problem(no, nc, M, 2 * no, seed=3, max_transfer=2)
    bst = bst_pq_state(p.bases, p.focks, p.ref_fock, max_q=mq, max_rank=mr, R=2, seed=4)
    tps = TPSCIstate.full_blocks(p.bases, p.focks)
    fo = flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops))
    fh = flatten_ham(p.clustered_ham)
    U = embedding(bst, p.bases, tps)
    Hb = U.T @ Oracle(fo, fh, tps).full_H() @ U
    return p, bst, Hb


@pytest.mark.parametrize("case", [(12, 3, 6, 2, 3), (8, 4, 4, 2, 2), (12, 3, 8, 3, 4)])
def test_bst_sigma_vs_projected_tps(eng, case):
    p, bst, Hb = _small(*case)
    sig = bst.copy_structure()
    build_sigma(sig, bst, p.cluster_ops, p.clustered_ham, engine=eng)
    print(case, len(bst), "rel", rel(sig.coeffs, Hb @ bst.coeffs), flush=True)
    assert rel(sig.coeffs, Hb @ bst.coeffs) < 1e-12
    # build_sigma accumulates like the reference's add!
    build_sigma(sig, bst, p.cluster_ops, p.clustered_ham, engine=eng)
    assert rel(sig.coeffs, 2 * Hb @ bst.coeffs) < 1e-12


def test_bst_identity_equals_tpsci_gpu(eng):
    p = build_problem(12, 3, 4, 24, seed=0)
    bi = bst_full_identity(p.bases, p.focks, R=2)
    bi.set_vector(np.random.Generator(np.random.PCG64(2)).normal(size=(len(bi), 2)))
    sig = bi.copy_structure()
    build_sigma(sig, bi, p.cluster_ops, p.clustered_ham, engine=eng)
    tps = TPSCIstate.full_blocks(p.bases, p.focks, R=2)
    tps.set_vector(bi.coeffs)
    s2 = tps_ci_matvec(tps, p.cluster_ops, p.clustered_ham, engine=eng)
    assert rel(sig.coeffs, s2) < 1e-12


def test_ci_solve_vs_dense(eng):
    p, bst, Hb = _small(12, 3, 6, 2, 3)
    w = np.linalg.eigvalsh(0.5 * (Hb + Hb.T))[:2]
    e, v = ci_solve(bst, p.cluster_ops, p.clustered_ham, conv_thresh=1e-9, max_iter=300,
                    engine=eng)
    print("ci_solve", e - w, v.davidson_stats, flush=True)
    assert np.abs(e - w).max() < 1e-10
    for prec in ("fp32", "bf16x2", "bf16"):
        e2, v2 = ci_solve(bst, p.cluster_ops, p.clustered_ham, conv_thresh=1e-8, max_iter=400,
                          precision=prec, engine=eng)
        print(prec, e2 - w, v2.davidson_stats, flush=True)
        assert np.abs(e2 - w).max() < 1e-6


def test_ci_solve_pseudo_canonical_precond(eng):
    # the reference's preconditioner (Hcmf diagonal in the pseudo-canonical
    # Tucker basis) and the plain diagonal of H reach the same energies; the
    # preconditioned result is returned in the rotated basis, where it is an
    # eigenvector of the projected H (checked through the embedding)
    from fermicg_amd.bst import form_1body_operator_diagonal
    p, bst, Hb = _small(12, 3, 8, 3, 4)
    w = np.linalg.eigvalsh(0.5 * (Hb + Hb.T))[:2]
    e1, v1 = ci_solve(bst, p.cluster_ops, p.clustered_ham, conv_thresh=1e-9, max_iter=300,
                      precond=True, engine=eng)
    e0, v0 = ci_solve(bst, p.cluster_ops, p.clustered_ham, conv_thresh=1e-9, max_iter=300,
                      precond=False, engine=eng)
    print("pseudo-canonical", v1.davidson_stats["iterations"], "plain",
          v0.davidson_stats["iterations"], "s2", v1.s2, flush=True)
    assert np.abs(e1 - w).max() < 1e-10 and np.abs(e0 - w).max() < 1e-10
    rot, _ = form_1body_operator_diagonal(bst, p.cluster_ops)
    assert rot.signature() == v1.signature()
    tps = TPSCIstate.full_blocks(p.bases, p.focks)
    U = embedding(rot, p.bases, tps)
    H = U.T @ Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                     flatten_ham(p.clustered_ham), tps).full_H() @ U
    r = H @ v1.coeffs - v1.coeffs * e1[None, :]
    assert np.abs(r).max() < 1e-8
    assert np.all(np.isfinite(v1.s2))


def test_bst_sigma_sampled_rows_four_clusters(eng):
    # 4 clusters of 4 orbitals, M=8, Tucker ranks <= 4: whole-space GPU sigma
    # against the C restatement of build_sigma! on 300 sampled bra blocks
    p, st = bst_medium(seed=11, n_clusters=4, max_roots=8, max_rank=4)
    sig = st.copy_structure()
    build_sigma(sig, st, p.cluster_ops, p.clustered_ham, engine=eng)
    orc = BstOracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                    flatten_ham(p.clustered_ham), st)
    rows = np.sort(np.random.Generator(np.random.PCG64(5)).choice(len(st.blocks), 300, replace=False))
    ref, idx = orc.sigma_rows(st.coeffs, rows)
    print("n_var", len(st), "rel", rel(sig.coeffs[idx], ref), flush=True)
    assert rel(sig.coeffs[idx], ref) < 1e-12
    # symmetry of H: d' H c == c' H d
    d = np.random.Generator(np.random.PCG64(6)).normal(size=st.coeffs.shape)
    sd = st.copy_structure()
    dd = st.copy_structure(); dd.set_vector(d)
    build_sigma(sd, dd, p.cluster_ops, p.clustered_ham, engine=eng)
    a = np.sum(d * sig.coeffs); b = np.sum(st.coeffs * sd.coeffs)
    assert abs(a - b) < 1e-11 * abs(a)

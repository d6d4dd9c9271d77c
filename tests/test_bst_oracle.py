"""CPU pins of the BST oracle (C restatement of build_sigma!):

* Tucker blocks with orthonormal factors: sigma_BST = U' H_TPS U c, with
  H_TPS from the TPSCI oracle (itself pinned by the FCI limit) and U the
  block-diagonal Kronecker embedding of the factors -- exact linear algebra
  that any dropped term, wrong overlap or transposed factor breaks;
* identity factors over whole Fock blocks reduce BST to TPSCI
  (reference test_bs.jl:6-34 compares BST with the dense-block BS path).
"""
import numpy as np
import pytest

from fermicg_amd.bst import bst_full_identity, bst_pq_state, embedding
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import cluster_sizes_of
from fermicg_amd.workloads import build_problem
from oracle.pyoracle import BstOracle, Oracle


@pytest.mark.parametrize("no,nc,M,mq,mr", [(12, 3, 6, 2, 3), (8, 4, 4, 2, 2)])
def test_bst_oracle_equals_projected_tps(oracle_lib, no, nc, M, mq, mr):
    p = build_problem(no, nc, M, 2 * no, seed=3, max_transfer=2)
    bst = bst_pq_state(p.bases, p.focks, p.ref_fock, max_q=mq, max_rank=mr, R=2, seed=4)
    tps = TPSCIstate.full_blocks(p.bases, p.focks)
    fo = flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops))
    fh = flatten_ham(p.clustered_ham)
    s = BstOracle(fo, fh, bst).sigma(bst.coeffs)
    U = embedding(bst, p.bases, tps)
    assert np.abs(U.T @ U - np.eye(U.shape[1])).max() < 1e-13
    ref = U.T @ (Oracle(fo, fh, tps).full_H() @ (U @ bst.coeffs))
    assert np.abs(s - ref).max() / np.abs(ref).max() < 1e-12


def test_bst_identity_equals_tpsci(oracle_lib):
    p = build_problem(12, 3, 4, 24, seed=0)
    bi = bst_full_identity(p.bases, p.focks)
    bi.set_vector(np.random.Generator(np.random.PCG64(1)).normal(size=(len(bi), 1)))
    fo = flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops))
    fh = flatten_ham(p.clustered_ham)
    s1 = BstOracle(fo, fh, bi).sigma(bi.coeffs)
    s2 = Oracle(fo, fh, TPSCIstate.full_blocks(p.bases, p.focks)).sigma(bi.coeffs)
    assert np.abs(s1 - s2).max() / np.abs(s2).max() < 1e-12

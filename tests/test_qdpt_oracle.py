"""Pins of the oracle's restatement of compute_qdpt_energy (CPU only).

* The reference's own case (test/test_qdpt.jl:40): Cr2 in five clusters,
  spin-adapted bases (M = 20), four roots in the spin-completed reference
  block solved in that 4-configuration space, QDPT at thresh_foi 1e-8:
  the four energies within the test's atol of 1e-12.
* What the mathematics fixes, by another route: with thresh_foi = 0 the
  first-order space holds every configuration some term reaches, so the
  QDPT matrix equals the Loewdin partitioning of the dense Hamiltonian over
  every Fock block, H_PP + H_PQ (E0[1] - diag H_QQ)^-1 H_QP, P the model
  configurations, Q all others (those no term reaches have H_QP = 0).  A
  wrong first-order set, a dropped term in the diagonal, the wrong root's
  E0 or a transposed H_ax all change these eigenvalues.  Bar: 1e-11.
"""
import numpy as np

from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import _full_space, first_order_focks
from oracle.pyoracle import Oracle, qdpt
from tests.pt2_cases import four_clusters
from tests.qdpt_cases import QDPT_REF, THRESH_FOI, Cr2


def _flat(ops, clusters, H):
    return flatten_ops(ops, [len(c) for c in clusters]), flatten_ham(H, 4)


def test_qdpt_cr2_reference_values(oracle_lib):
    p = Cr2()
    fops, fham = _flat(p.ops, p.clusters, p.H)
    st = p.start()
    w, v = np.linalg.eigh(Oracle(fops, fham, st).full_H())
    st.set_vector(v[:, :4])
    target = _full_space(p.ops, first_order_focks(st, p.ops, p.H), 4)
    e, nx = qdpt(fops, fham, st, target, THRESH_FOI, nthreads=8)
    print("first-order configurations", nx, "dE", e - QDPT_REF, flush=True)
    assert len(e) == 4
    assert np.abs(e - QDPT_REF).max() < 1e-12


def test_qdpt_matches_dense_partitioning(oracle_lib):
    case = four_clusters(seed=6, R=2)
    st = case.state
    fops, fham = _flat(case.ops, case.clusters, case.H)
    target = _full_space(case.ops, first_order_focks(st, case.ops, case.H), st.R)
    e, nx = qdpt(fops, fham, st, target, 0.0)
    full = TPSCIstate.full_blocks(case.bases, case.all_focks, R=1)
    H = Oracle(fops, fham, full).full_H()
    focks = full.focks()
    index = {(focks[fi], tuple(int(x) for x in full.cfg[i])): i
             for fi, a, b in full.blocks() for i in range(a, b)}
    sf = st.focks()
    P = np.array([index[(sf[fi], tuple(int(x) for x in st.cfg[i]))]
                  for fi, a, b in st.blocks() for i in range(a, b)])
    Q = np.setdiff1d(np.arange(len(full)), P)
    c = st.coeffs
    e0 = np.einsum("ir,ij,jr->r", c, H[np.ix_(P, P)], c)
    heff = H[np.ix_(P, P)] + (H[np.ix_(P, Q)] / (e0[0] - np.diag(H)[Q])[None, :]) @ H[np.ix_(Q, P)]
    ref = np.linalg.eigvalsh(0.5 * (heff + heff.T))
    print("model", len(P), "first-order", nx, "of", len(Q), flush=True)
    assert 0 < nx < len(Q) and len(e) == len(P)
    assert np.abs(e - ref).max() < 1e-11 * np.abs(ref).max()

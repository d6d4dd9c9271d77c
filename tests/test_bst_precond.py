"""form_1body_operator_diagonal (src/tucker_pt.jl:284-355) on the CPU: the
pseudo-canonical rotation leaves the BST state unchanged (same vector once
embedded in the full tensor-product space), keeps every factor orthonormal,
and the returned diagonal is the diagonal of sum_c Hcmf_c in the rotated
basis, computed independently through the embedding."""
import numpy as np

from fermicg_amd.bst import bst_pq_state, embedding, form_1body_operator_diagonal
from fermicg_amd.state import TPSCIstate
from fermicg_amd.workloads import build_problem


def _h0_full(p, tps):
    """sum_c Hcmf_c as a dense matrix over the full-block TPS space."""
    n = len(tps)
    H0 = np.zeros((n, n))
    N = len(p.bases)
    for fi, a, b in tps.blocks():
        f = tps.focks()[fi]
        M = [p.bases[c].n_states(f[c]) for c in range(N)]
        blk = np.zeros((b - a, b - a))
        for c in range(N):
            K = np.ones((1, 1))
            for c2 in range(N):
                m = np.asarray(p.cluster_ops[c]["Hcmf"][(f[c], f[c])]) if c2 == c else np.eye(M[c2])
                K = np.kron(m, K)
            blk += K
        H0[a:b, a:b] = blk
    return H0


def test_pseudo_canonical_rotation():
    p = build_problem(12, 3, 6, 24, seed=3, max_transfer=2)
    st = bst_pq_state(p.bases, p.focks, p.ref_fock, max_q=2, max_rank=3, R=2, seed=4)
    tps = TPSCIstate.full_blocks(p.bases, p.focks)
    rot, hd = form_1body_operator_diagonal(st, p.cluster_ops)
    U0, U1 = embedding(st, p.bases, tps), embedding(rot, p.bases, tps)
    assert np.abs(U0 @ st.coeffs - U1 @ rot.coeffs).max() < 1e-13
    assert np.abs(U1.T @ U1 - np.eye(U1.shape[1])).max() < 1e-12
    H0 = U1.T @ _h0_full(p, tps) @ U1
    assert np.abs(np.diag(H0) - hd).max() < 1e-12
    # within every block, each active cluster's H0 is now diagonal
    off = rot.var_offsets()
    for b in range(len(rot.blocks)):
        blk = H0[off[b]:off[b + 1], off[b]:off[b + 1]]
        assert np.abs(blk - np.diag(np.diag(blk))).max() < 1e-10
    # without the rotation: the same state and the diagonal of U' H0 U
    same, hd0 = form_1body_operator_diagonal(st, p.cluster_ops, pseudo_canon=False)
    assert np.array_equal(same.coeffs, st.coeffs)
    assert np.abs(np.diag(U0.T @ _h0_full(p, tps) @ U0) - hd0).max() < 1e-12

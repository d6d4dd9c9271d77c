"""A CPU engine for the TPSCI drivers of fermicg_amd.tpsci, backed by the C
oracle (test infrastructure only).

fermicg_amd.tpsci's drivers (tpsci_ci, compute_pt2_energy,
compute_pt1_wavefunction, ...) take an `engine=`; the product engine runs
every sigma, diagonal, solve, open matvec and PT2 in libfcg_hip.so.  This
one answers the same calls from the oracle's restatements of the
reference (dense H from build_full_H, the open matvec with the
_collect_significant_thread! element test, compute_pt2_energy), so the
drivers' own logic and the oracle's screening rule can be checked on the
CPU against values the reference's tests hold.
"""
import numpy as np

from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.tpsci import cluster_sizes_of
from oracle.pyoracle import Oracle


class _DensePlan:
    """sigma / diagonal / eigenpairs of H over one space, from the oracle's
    dense build_full_H."""

    def __init__(self, orc):
        self.H = orc.full_H()
        self.nvar = self.nvar_ket = self.H.shape[0]

    def sigma(self, c, precision="fp64"):
        return self.H @ np.asarray(c, np.float64)

    def diagonal(self):
        return np.diag(self.H).copy()

    def davidson(self, v0, conv_thresh=1e-8, max_iter=200, max_ss_vecs=12, lindep_thresh=1e-12,
                 precision="fp64", refresh_k=0, precond=True):
        v0 = np.asarray(v0, np.float64)
        R = 1 if v0.ndim == 1 else v0.shape[1]
        w, v = np.linalg.eigh(self.H)
        return w[:R].copy(), v[:, :R].copy(), {"iterations": 0, "sigma_fp64": 0,
                                               "sigma_low": 0, "converged": True,
                                               "residual": [0.0] * R}


class OracleEngine:
    """Drop-in for fermicg_amd.tpsci.Engine on the CPU (oracle-backed)."""

    def __init__(self, nthreads=0):
        self.nthreads = nthreads
        self._plans = {}
        self._local = {}
        self._flat = {}

    def _flat_inputs(self, cluster_ops, clustered_ham, nbody=4):
        key = (id(cluster_ops), id(clustered_ham), nbody)
        hit = self._flat.get(key)
        if hit is None or hit[0] is not cluster_ops or hit[1] is not clustered_ham:
            hit = (cluster_ops, clustered_ham,
                   flatten_ops(cluster_ops, cluster_sizes_of(cluster_ops)),
                   flatten_ham(clustered_ham, nbody))
            self._flat[key] = hit
        return hit[2], hit[3]

    def oracle(self, state, cluster_ops, clustered_ham, nbody=4):
        fo, fh = self._flat_inputs(cluster_ops, clustered_ham, nbody)
        return Oracle(fo, fh, state)

    def tpsci_plan(self, state, cluster_ops, clustered_ham, nbody=4, precisions=("fp64",)):
        sig = (state.fock.tobytes(), state.fock_start.tobytes(), state.cfg.tobytes())
        key = (id(cluster_ops), id(clustered_ham), nbody, hash(sig))
        pl = self._plans.get(key)
        if pl is None:
            pl = _DensePlan(self.oracle(state, cluster_ops, clustered_ham, nbody))
            self._plans[key] = pl
        return pl

    def space_diagonal(self, state, cluster_ops, clustered_ham):
        return self.oracle(state, cluster_ops, clustered_ham).diagonal()

    def local_operator(self, clustered_ham, op_string):
        from fermicg_amd.operator import extract_1body_operator
        key = (id(clustered_ham), op_string)
        hit = self._local.get(key)
        if hit is None or hit[0] is not clustered_ham:
            hit = (clustered_ham, extract_1body_operator(clustered_ham, op_string=op_string))
            self._local[key] = hit
        return hit[1]

    def s2_operator(self, cluster_ops):
        from fermicg_amd.operator import extract_S2
        sizes = tuple(cluster_sizes_of(cluster_ops))
        key = ("S2", sizes)
        hit = self._local.get(key)
        if hit is None:
            hit = (None, extract_S2(len(sizes), sizes))
            self._local[key] = hit
        return hit[1]

    def open_sigma(self, ket_state, target_state, cluster_ops, clustered_ham, nbody=4,
                   thresh=0.0):
        orc = self.oracle(ket_state, cluster_ops, clustered_ham, nbody)
        sizes = np.diff(target_state.fock_start)
        return orc.open_sigma(ket_state.coeffs, target_state.fock, sizes, nbody=nbody,
                              thresh=thresh, nthreads=self.nthreads)

    def pt2(self, ci_vector, cluster_ops, clustered_ham, E0, nbody=4, H0="Hcmf", thresh=0.0,
            method="dense"):
        from fermicg_amd.cluster_ops import OP_NAMES
        orc = self.oracle(ci_vector, cluster_ops, clustered_ham, nbody)
        e2, n_jobs = orc.pt2(ci_vector.coeffs, E0, nbody=nbody, h0_op=OP_NAMES.index(H0),
                             thresh=thresh, nthreads=self.nthreads)
        return e2, {"n_target_fock": n_jobs, "n_first_order": -1, "seconds": 0.0}

    def bst_pair_sigma(self, bra, ket, cluster_ops, clustered_ham, nbody=4):
        """H c_ket on bra's blocks: the C restatement of build_sigma! over the
        union of both block lists, ket coefficients placed after zeros."""
        from fermicg_amd.bst import BSTstate
        from oracle.pyoracle import BstOracle
        u = BSTstate(bra.N, ket.coeffs.shape[1])
        for st in (bra, ket):
            for fi, rgs, facs in st.blocks:
                u.add_block(st.focks[fi], rgs, facs)
        n_bra = len(bra)
        u.coeffs = np.vstack([np.zeros((n_bra, ket.coeffs.shape[1])), ket.coeffs])
        fo, fh = self._flat_inputs(cluster_ops, clustered_ham, nbody)
        orc = BstOracle(fo, fh, u)
        n_bra_boxes = len(bra.flat().box_fock)
        s, idx = orc.sigma_rows(u.coeffs, np.arange(n_bra_boxes), nthreads=self.nthreads)
        assert len(idx) == n_bra
        return s

    def bst_ci_solve(self, vec, cluster_ops, clustered_ham, conv_thresh=1e-5, max_iter=40,
                     max_ss_vecs=12, lindep_thresh=1e-10, nbody=4):
        """ci_solve's result from the dense H of the space (columns from the
        restated sigma), in the pseudo-canonical basis ci_solve returns."""
        from fermicg_amd.bst import form_1body_operator_diagonal
        rot, _ = form_1body_operator_diagonal(vec, cluster_ops, H0="Hcmf", pseudo_canon=True)
        n, R = len(rot), rot.coeffs.shape[1]
        H = np.zeros((n, n))
        probe = rot.copy_structure(R=1)
        for j0 in range(0, n, 64):
            j1 = min(n, j0 + 64)
            probe.R = j1 - j0
            probe.coeffs = np.zeros((n, j1 - j0))
            probe.coeffs[np.arange(j0, j1), np.arange(j1 - j0)] = 1.0
            H[:, j0:j1] = self.bst_pair_sigma(rot, probe, cluster_ops, clustered_ham, nbody)
        w, v = np.linalg.eigh(0.5 * (H + H.T))
        out = rot.copy_structure(R=R)
        out.set_vector(v[:, :R])
        return w[:R].copy(), out

    def drop_plans(self, keep=()):
        for k in [k for k in self._plans if k not in keep]:
            del self._plans[k]

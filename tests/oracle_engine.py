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

    def drop_plans(self, keep=()):
        for k in [k for k in self._plans if k not in keep]:
            del self._plans[k]

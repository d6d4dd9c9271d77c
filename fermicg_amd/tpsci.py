"""TPSCI entry points on the GPU, mirroring the reference's Julia API.

    tps_ci_matvec(ci_vector, cluster_ops, clustered_ham)      tpsci_outer.jl:462
    compute_diagonal(ci_vector, cluster_ops, clustered_ham)   tpsci_outer.jl:694
    tps_ci_davidson(ci_vector, cluster_ops, clustered_ham; …) tpsci_outer.jl:364

Same arguments, same meaning, same output layout (rows in the TPSCIstate's
flattening order, one column per root).  All arithmetic runs in
libfcg_hip.so; device-resident copies of ClusterOps, the ClusteredOperator
and the Hamiltonian blocks of each space are cached on an Engine and reused
across calls (the reference rebuilds nothing between Davidson iterations
either; its op cache is term.cache, tucker_inner.jl:90-159).
"""
import numpy as np

from . import _lib
from .marshal import flatten_ham, flatten_ops


def cluster_sizes_of(cluster_ops):
    """Orbitals per cluster, read from the single-creation Γ blocks."""
    sizes = []
    for ops in cluster_ops:
        blk = next(iter(ops["A"].values()), None) if "A" in ops else None
        if blk is None:
            blk = next(iter(ops["B"].values()))
        sizes.append(blk.shape[0])
    return sizes


class Engine:
    """Device context plus caches of uploaded inputs and plans."""

    def __init__(self, device=0):
        self.ctx = _lib.Context(device)
        self._ops = {}
        self._ham = {}
        self._plans = {}

    def ops(self, cluster_ops):
        key = id(cluster_ops)
        hit = self._ops.get(key)
        if hit is None or hit[0] is not cluster_ops:
            fops = flatten_ops(cluster_ops, cluster_sizes_of(cluster_ops))
            hit = (cluster_ops, _lib.DeviceOps.upload(self.ctx, fops), fops)
            self._ops[key] = hit
        return hit[1]

    def ham(self, clustered_ham, nbody=4):
        key = (id(clustered_ham), nbody)
        hit = self._ham.get(key)
        if hit is None or hit[0] is not clustered_ham:
            fh = flatten_ham(clustered_ham, nbody)
            hit = (clustered_ham, _lib.DeviceHam.upload(self.ctx, fh, clustered_ham.n_clusters))
            self._ham[key] = hit
        return hit[1]

    def tpsci_plan(self, state, cluster_ops, clustered_ham, nbody=4, precisions=("fp64",)):
        sig = (state.fock.tobytes(), state.fock_start.tobytes(), state.cfg.tobytes())
        key = (id(cluster_ops), id(clustered_ham), nbody, hash(sig), tuple(sorted(precisions)))
        pl = self._plans.get(key)
        if pl is None:
            sp = _lib.DeviceSpace.tpsci(self.ctx, state)
            pl = _lib.Plan.create(self.ctx, self.ham(clustered_ham, nbody), self.ops(cluster_ops),
                                  sp, nbody=nbody, precisions=precisions)
            self._plans[key] = pl
        return pl

    def clear(self):
        self._plans.clear()
        self._ham.clear()
        self._ops.clear()


_ENGINE = None


def default_engine():
    global _ENGINE
    if _ENGINE is None:
        _ENGINE = Engine(0)
    return _ENGINE


def tps_ci_matvec(ci_vector, cluster_ops, clustered_ham, precision="fp64", engine=None):
    """sigma = H c over the configurations of `ci_vector` (n_cfg x R)."""
    eng = engine or default_engine()
    pl = eng.tpsci_plan(ci_vector, cluster_ops, clustered_ham,
                        precisions=tuple(sorted({"fp64", precision})))
    return pl.sigma(ci_vector.coeffs, precision)


def compute_diagonal(ci_vector, cluster_ops, clustered_ham, engine=None):
    eng = engine or default_engine()
    return eng.tpsci_plan(ci_vector, cluster_ops, clustered_ham).diagonal()


def tps_ci_davidson(ci_vector, cluster_ops, clustered_ham, conv_thresh=1e-5, lindep_thresh=1e-12,
                    max_ss_vecs=12, max_iter=40, precond=True, precision="fp64", refresh_k=4,
                    engine=None):
    """Lowest R eigenpairs in the space of `ci_vector`, starting from its
    coefficients.  Returns (energies, TPSCIstate with the eigenvectors)."""
    eng = engine or default_engine()
    pl = eng.tpsci_plan(ci_vector, cluster_ops, clustered_ham,
                        precisions=tuple(sorted({"fp64", precision})))
    e, v, stats = pl.davidson(ci_vector.coeffs, conv_thresh=conv_thresh, max_iter=max_iter,
                              max_ss_vecs=max_ss_vecs, lindep_thresh=lindep_thresh,
                              precision=precision, refresh_k=refresh_k, precond=precond)
    out = _copy_state(ci_vector)
    out.set_vector(v)
    out.davidson_stats = stats
    return e, out


def _copy_state(s):
    from .state import TPSCIstate
    o = TPSCIstate(s.N, s.R)
    o.fock = s.fock.copy()
    o.fock_start = s.fock_start.copy()
    o.cfg = s.cfg.copy()
    o.coeffs = s.coeffs.copy()
    return o

"""Seeded synthetic workloads (the benchmark configurations).

tpsci_small   12 orbitals, 12 electrons, 3 clusters of 4 orbitals, lowest 4
              states per Fock sector, every ClusterConfig of every Fock
              configuration with at most 2 electrons transferred from
              (2,2)x3: 73 Fock blocks, 4,672 configurations; 1 root.
bst_medium    24 orbitals, 6 clusters of 4, 16 states per sector, Tucker
              blocks P=1:1, Q=2:M with at most 2 clusters in Q, random ranks
              1..min(8,|Q|) and orthonormal factors, cores N(0,1)*10^-transfer,
              over the 1,441 Fock configurations with <= 2 transferred; 2 roots.
Integrals: fermicg_amd.integrals.synthetic_ints (seed, L = 2 x orbitals).
"""
from dataclasses import dataclass, field

import numpy as np

from .cluster_ops import add_cmf_operators, compute_cluster_eigenbasis, compute_cluster_ops, product_state_rdm1
from .integrals import synthetic_ints
from .operator import extract_ClusteredTerms
from .state import TPSCIstate, fock_configs_within, n_transferred


@dataclass
class Problem:
    ints: object
    clusters: list
    ref_fock: tuple
    bases: list
    cluster_ops: list
    clustered_ham: object
    focks: list
    meta: dict = field(default_factory=dict)


def build_problem(n_orb, n_clusters, max_roots, n_aux, seed, max_transfer=2, cmf=True):
    per = n_orb // n_clusters
    clusters = [list(range(per * i, per * (i + 1))) for i in range(n_clusters)]
    half = per // 2
    ref = tuple((half, half) for _ in range(n_clusters))
    ints = synthetic_ints(n_orb, n_aux, seed)
    bases = compute_cluster_eigenbasis(ints, clusters, max_roots=max_roots)
    ops = compute_cluster_ops(bases, ints)
    if cmf:
        Da, Db = product_state_rdm1(bases, n_orb, ref)
        add_cmf_operators(ops, bases, ints, Da, Db)
    H = extract_ClusteredTerms(ints, clusters)
    focks = fock_configs_within(bases, ref, max_transfer)
    return Problem(ints, clusters, ref, bases, ops, H, focks)


def tpsci_small(seed=0, R=1):
    """12 orbitals / 3 clusters / M=4 / <=2 transferred, full Fock blocks."""
    p = build_problem(12, 3, 4, 24, seed)
    st = TPSCIstate.full_blocks(p.bases, p.focks, R=R)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    v = rng.normal(size=(len(st), R))
    v /= np.linalg.norm(v, axis=0)
    st.set_vector(v)
    p.meta = {"workload": "tpsci_12o_3c_M4", "n_fock": len(p.focks), "n_cfg": len(st)}
    return p, st


def bst_medium(seed=1, R=2, n_clusters=6, max_roots=16, max_rank=8):
    """24 orbitals / 6 clusters / M=16 / Tucker P-Q blocks with <= 2 Q
    clusters and ranks <= 8 over the Fock configurations with <= 2
    electrons transferred."""
    from .bst import bst_pq_state
    p = build_problem(4 * n_clusters, n_clusters, max_roots, 8 * n_clusters, seed)
    st = bst_pq_state(p.bases, p.focks, p.ref_fock, max_q=2, max_rank=max_rank, R=R,
                      seed=seed + 2000)
    p.meta = {"workload": "bst_%do_%dc_M%d_r%d" % (4 * n_clusters, n_clusters, max_roots, max_rank),
              "n_fock": len(p.focks), "n_blocks": len(st.blocks), "n_var": len(st)}
    return p, st

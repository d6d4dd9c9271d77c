"""Seeded PT2 test cases shared by the CPU oracle pins and the GPU parity tests.

Each case is a variational TPSCI state (a seeded subset of the
configurations of the Fock blocks near the reference, holding the lowest R
eigenvectors of H in that space) on a synthetic problem, plus the problem's
ClusterOps and ClusteredOperator.
"""
import numpy as np

from fermicg_amd.cluster_ops import (add_cmf_operators, compute_cluster_eigenbasis,
                                     compute_cluster_ops, product_state_rdm1)
from fermicg_amd.integrals import synthetic_ints
from fermicg_amd.operator import extract_ClusteredTerms
from fermicg_amd.state import TPSCIstate, fock_configs_within
from fermicg_amd.tpsci import orthonormalize


class Case:
    pass


def make_case(n_orb, n_clusters, ref, max_roots, seed, R=2, var_transfer=1, keep=0.5,
              full_blocks=False):
    per = n_orb // n_clusters
    clusters = [list(range(per * i, per * (i + 1))) for i in range(n_clusters)]
    ints = synthetic_ints(n_orb, 2 * n_orb, seed)
    bases = compute_cluster_eigenbasis(ints, clusters, max_roots=max_roots)
    ops = compute_cluster_ops(bases, ints)
    Da, Db = product_state_rdm1(bases, n_orb, ref)
    add_cmf_operators(ops, bases, ints, Da, Db)
    H = extract_ClusteredTerms(ints, clusters)
    var_focks = fock_configs_within(bases, ref, var_transfer)
    full = TPSCIstate.full_blocks(bases, var_focks, R=R)
    rng = np.random.Generator(np.random.PCG64(seed + 77))
    if full_blocks:
        st = full
    else:
        mask = rng.random(len(full)) < keep
        mask[0] = True
        blocks = []
        for fi, a, b in full.blocks():
            confs = [tuple(int(x) for x in full.cfg[i]) for i in range(a, b) if mask[i]]
            if confs:
                blocks.append((full.focks()[fi], confs))
        st = TPSCIstate.from_blocks(n_clusters, blocks, R=R)
    # lowest R eigenvectors of H in the variational space (dense, oracle)
    from fermicg_amd.marshal import flatten_ham, flatten_ops
    from oracle.pyoracle import Oracle
    Hv = Oracle(flatten_ops(ops, [len(x) for x in clusters]), flatten_ham(H), st).full_H()
    w, U = np.linalg.eigh(Hv)
    st.set_vector(orthonormalize(U[:, :R]))
    c = Case()
    c.ints, c.clusters, c.ref, c.bases, c.ops, c.H, c.state = ints, clusters, ref, bases, ops, H, st
    c.all_focks = fock_configs_within(bases, ref, 100)
    return c


def four_clusters(seed=3, R=2):
    """8 orbitals in 4 clusters of 2, 2 alpha + 2 beta electrons (3- and
    4-cluster terms present); complete cluster bases."""
    return make_case(8, 4, ((1, 0), (0, 1), (1, 0), (0, 1)), 16, seed, R=R)


def three_clusters_m4(seed=0, R=2, full_blocks=False):
    """12 orbitals in 3 clusters of 4 at half filling, lowest 4 states per
    sector (the TPSCI benchmark problem), variational blocks within one
    transferred electron."""
    return make_case(12, 3, ((2, 2), (2, 2), (2, 2)), 4, seed, R=R, full_blocks=full_blocks)


def six_clusters_rank1(seed=21, R=2, frac=0.6):
    """24 orbitals in 6 clusters of 4 at half filling, lowest 4 states per
    sector, a hash-selected rank <= 1 variational space over the Fock blocks
    within one transferred electron (710 configurations), and an orthonormal
    state with a dominant reference configuration and small seeded
    coefficients elsewhere (the shape screening is made for)."""
    from fermicg_amd.workloads import build_problem, rank1_space
    p = build_problem(24, 6, 4, 48, seed=seed, max_transfer=1)
    st = rank1_space(p, R=R, frac=frac, seed=seed)
    rng = np.random.default_rng(seed + 5)
    v = rng.normal(size=(len(st), R)) * 0.05
    for r in range(R):
        v[r, r] += 1.0
    st.set_vector(orthonormalize(v))
    c = Case()
    c.ints, c.clusters, c.ref, c.bases = p.ints, p.clusters, p.ref_fock, p.bases
    c.ops, c.H, c.state = p.cluster_ops, p.clustered_ham, st
    return c


def twelve_clusters_rank1(M=4, nnz=6, seed=4, frac=0.05):
    """The 48-orbital, 12-cluster problem (clusters of 4 at half filling,
    lowest M states per sector, Fock blocks within 2 transferred electrons)
    with a hash-selected rank <= 1 variational space and a state whose
    coefficients are zero except on the reference configuration (1) and
    nnz - 1 seeded configurations (N(0, 0.05)), normalised: the shape of the
    12-cluster first-order workload, small enough for the oracle."""
    from fermicg_amd.workloads import build_problem, rank1_space
    p = build_problem(48, 12, M, 96, seed, max_transfer=2)
    st = rank1_space(p, R=1, frac=frac, seed=seed)
    ref_row = int(st.fock_start[p.focks.index(p.ref_fock)])
    rng = np.random.default_rng(seed + 9)
    v = np.zeros((len(st), 1))
    v[ref_row, 0] = 1.0
    pick = rng.choice(len(st), size=nnz - 1, replace=False)
    v[pick, 0] += rng.normal(size=nnz - 1) * 0.05
    st.set_vector(v / np.linalg.norm(v))
    c = Case()
    c.ints, c.clusters, c.ref, c.bases = p.ints, p.clusters, p.ref_fock, p.bases
    c.ops, c.H, c.state = p.cluster_ops, p.clustered_ham, st
    return c

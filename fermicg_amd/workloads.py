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


def pt2_medium(seed=11, M=7, R=2):
    """PT2 workload: 16 orbitals / 4 clusters of 4 at half filling / M states
    per sector; variational space = every configuration of the Fock blocks
    within 1 transferred electron (25 blocks, M=7: 60,025 configurations),
    R orthonormal seeded vectors.  The first-order space is every
    configuration of the Fock blocks reached by one TransferConfig."""
    from .tpsci import orthonormalize
    p = build_problem(16, 4, M, 32, seed, max_transfer=1)
    st = TPSCIstate.full_blocks(p.bases, p.focks, R=R)
    rng = np.random.Generator(np.random.PCG64(seed + 3000))
    st.set_vector(orthonormalize(rng.normal(size=(len(st), R))))
    p.meta = {"workload": "pt2_16o_4c_M%d" % M, "n_fock": len(p.focks), "n_cfg": len(st)}
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


def splitmix64(x):
    """splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def rank1_space(p, R=4, frac=0.159, seed=2):
    """Rank <= 1 ClusterConfigs over the problem's Fock configurations: in
    every Fock block the all-ground configuration (1,...,1) and each
    configuration with exactly one cluster c in state s >= 2, the latter kept
    when splitmix64(seed ^ key) < frac * 2^64 with key = (Fock index << 16) |
    (c << 8) | s.  Fock blocks in the problem's order; inside a block the
    ground configuration first, then by cluster, then by state."""
    N = len(p.bases)
    focks = p.focks
    M = np.array([[p.bases[c].n_states(f[c]) for c in range(N)] for f in focks], np.int64)
    nf = len(focks)
    # candidate (fock, cluster, state) triples, s = 2..M
    fi, ci, si = [], [], []
    for c in range(N):
        for s in range(2, int(M[:, c].max()) + 1):
            f = np.nonzero(M[:, c] >= s)[0]
            fi.append(f); ci.append(np.full(len(f), c)); si.append(np.full(len(f), s))
    fi = np.concatenate(fi); ci = np.concatenate(ci); si = np.concatenate(si)
    key = (fi.astype(np.uint64) << np.uint64(16)) | (ci.astype(np.uint64) << np.uint64(8)) | si.astype(np.uint64)
    h = splitmix64(key ^ np.uint64(seed))
    keep = h < np.uint64(min(int(frac * 2.0 ** 64), 2 ** 64 - 1))
    if frac >= 1.0:
        keep[:] = True
    fi, ci, si = fi[keep], ci[keep], si[keep]
    order = np.lexsort((si, ci, fi))
    fi, ci, si = fi[order], ci[order], si[order]
    counts = np.bincount(fi, minlength=nf) + 1          # + the ground configuration
    start = np.zeros(nf + 1, np.int64)
    start[1:] = np.cumsum(counts)
    cfg = np.ones((start[-1], N), np.int16)
    # row of each excited configuration: after its block's ground row
    rows = start[fi] + 1 + (np.arange(len(fi)) - np.searchsorted(fi, fi))
    cfg[rows, ci] = si
    from .state import TPSCIstate
    st = TPSCIstate(N, R)
    st.fock = np.array(focks, np.int16).reshape(nf, N, 2)
    st.fock_start = start
    st.cfg = cfg
    st.coeffs = np.zeros((len(cfg), R))
    return st


def tpsci_sparse(seed=2, R=4, n_clusters=8, max_roots=16, max_transfer=3, frac=0.159):
    """32 orbitals / 8 clusters / M=16 / Fock configurations within 3
    transferred electrons, rank <= 1 configurations selected by hash (about
    2e6); R random normalised start vectors."""
    p = build_problem(4 * n_clusters, n_clusters, max_roots, 8 * n_clusters, seed,
                      max_transfer=max_transfer)
    st = rank1_space(p, R=R, frac=frac, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed + 3000))
    v = rng.normal(size=(len(st), R))
    v /= np.linalg.norm(v, axis=0)
    st.set_vector(v)
    p.meta = {"workload": "tpsci_%do_%dc_M%d_rank1" % (4 * n_clusters, n_clusters, max_roots),
              "n_fock": len(p.focks), "n_cfg": len(st)}
    return p, st


def pt2_12c_M32(seed=4, R=1, M=32):
    """48 orbitals / 12 clusters of 4 at half filling / M states per sector;
    variational space = every rank <= 1 configuration of the 26,533 Fock
    configurations within 2 transferred electrons (9,020,329 at M=32).  The
    state is built on the GPU by cmf_first_order_state."""
    p = build_problem(48, 12, M, 96, seed, max_transfer=2)
    st = rank1_space(p, R=R, frac=1.0, seed=seed)
    p.meta = {"workload": "pt2_48o_12c_M%d_rank1" % M, "n_fock": len(p.focks), "n_cfg": len(st)}
    return p, st


def cmf_first_order_state(p, st, engine, shift=1.0, select=0.0):
    """The CMF reference (the ground configuration of the reference Fock
    block) plus a regularised first-order correction on the space: every
    other configuration J gets a_J = -(H|ref>)_J / (|H0_JJ - H0_ref| + shift),
    with H0 the clusters' Hcmf and shift in Eh (synthetic integrals put some
    charge-transfer blocks below the reference in H0, so the plain
    denominators change sign and come near zero).  Amplitudes with
    |a_J| <= select are set to zero (one TPSCI selection step from the
    reference; the configurations stay in the space).  H|ref> is the
    matrix-free sigma of the rank <= 1 space applied to the reference alone.
    Normalised; returns the coefficient vector (n x 1)."""
    ref_f = p.focks.index(p.ref_fock)
    ref_row = int(st.fock_start[ref_f])
    assert all(int(x) == 1 for x in st.cfg[ref_row])
    x = np.zeros((len(st), 1))
    x[ref_row, 0] = 1.0
    pl = engine.tpsci_plan(st, p.cluster_ops, p.clustered_ham, matrix_free=True)
    sig = pl.sigma(x)[:, 0]
    h0 = engine.local_operator(p.clustered_ham, "Hcmf")
    hd0 = engine.tpsci_plan(st, p.cluster_ops, h0, matrix_free=True).diagonal()
    c = -sig / (np.abs(hd0 - hd0[ref_row]) + shift)
    c[np.abs(c) <= select] = 0.0
    c[ref_row] = 1.0
    c /= np.linalg.norm(c)
    return c[:, None]

"""Flattening of ClusterOps / ClusteredOperator / TPSCIstate into the plain
arrays that cross the C ABI (the on-host wire format).

Sector code of a cluster Fock sector (nα, nβ): na * S + nb with
S = max cluster size + 1.  Γ blocks are addressed through a dense table
gblock[cluster][op][code_bra][code_ket] -> block id (or -1), so the device
never hashes strings or tuples.  Γ data are concatenated column-major as in
ClusterOps (n^k, M_bra, M_ket).  Terms are ordered by TransferConfig group;
tc_start[g]..tc_start[g+1] are the terms of group g.  All integers are
little-endian int32 / int16 / int64 as noted.
"""
from dataclasses import dataclass

import numpy as np

from .cluster_ops import OP_ID, OP_NAMES


@dataclass
class FlatOps:
    n_clusters: int
    S: int
    n_ops: int
    gblock: np.ndarray   # int32 [n_clusters, n_ops, S*S, S*S]
    gb_off: np.ndarray   # int64 [n_blocks]
    gb_dims: np.ndarray  # int32 [n_blocks, 3] (np, m_bra, m_ket)
    gdata: np.ndarray    # float64 concatenated blocks
    nstates: np.ndarray  # int32 [n_clusters, S*S] states per sector (0 = absent)
    cluster_size: np.ndarray  # int32 [n_clusters]


@dataclass
class FlatHam:
    n_terms: int
    nbody: np.ndarray    # int32 [T]
    clusters: np.ndarray # int32 [T,4] (-1 pad)
    ops: np.ndarray      # int32 [T,4]
    parity: np.ndarray   # int32 [T,4]
    dims: np.ndarray     # int32 [T,4] ints dims (1 pad)
    delta: np.ndarray    # int16 [T,4,2]
    ioff: np.ndarray     # int64 [T]
    ints: np.ndarray     # float64
    n_tc: int
    tc: np.ndarray       # int16 [n_tc, N, 2]
    tc_start: np.ndarray # int32 [n_tc+1]
    terms: list          # python ClusteredTerm objects in flat order


def flatten_ops(all_ops, cluster_sizes):
    n_clusters = len(all_ops)
    S = int(max(cluster_sizes)) + 1
    n_ops = len(OP_NAMES)
    gblock = -np.ones((n_clusters, n_ops, S * S, S * S), dtype=np.int32)
    nstates = np.zeros((n_clusters, S * S), dtype=np.int32)
    offs, dims, chunks = [], [], []
    at = 0
    for c, ops in enumerate(all_ops):
        for (sb, sk), G in ops["H"].items():
            nstates[c, sb[0] * S + sb[1]] = G.shape[0]
        for name, blocks in ops.items():
            o = OP_ID[name]
            for (sb, sk), G in blocks.items():
                if G.ndim == 2:
                    npd, mb, mk = 1, G.shape[0], G.shape[1]
                else:
                    npd, mb, mk = G.shape
                gblock[c, o, sb[0] * S + sb[1], sk[0] * S + sk[1]] = len(offs)
                offs.append(at)
                dims.append((npd, mb, mk))
                flat = np.asarray(G, dtype=np.float64).ravel(order="F")
                chunks.append(flat)
                at += flat.size
    gdata = np.concatenate(chunks) if chunks else np.zeros(0)
    return FlatOps(n_clusters, S, n_ops, gblock, np.array(offs, dtype=np.int64),
                   np.array(dims, dtype=np.int32).reshape(-1, 3), gdata, nstates,
                   np.asarray(cluster_sizes, dtype=np.int32))


def flatten_ham(H, nbody_max=4):
    N = H.n_clusters
    tcs, tc_start, terms = [], [0], []
    for tc, ts in H:
        kept = [t for t in ts if t.nbody <= nbody_max]
        if not kept:
            continue
        tcs.append(tc)
        terms.extend(kept)
        tc_start.append(len(terms))
    T = len(terms)
    nbody = np.zeros(T, np.int32)
    clusters = -np.ones((T, 4), np.int32)
    ops = -np.ones((T, 4), np.int32)
    parity = np.zeros((T, 4), np.int32)
    dims = np.ones((T, 4), np.int32)
    delta = np.zeros((T, 4, 2), np.int16)
    ioff = np.zeros(T, np.int64)
    chunks, at = [], 0
    for i, t in enumerate(terms):
        k = t.nbody
        nbody[i] = k
        clusters[i, :k] = t.clusters
        ops[i, :k] = [OP_ID[o] for o in t.ops]
        parity[i, :k] = t.parity
        delta[i, :k] = t.delta
        v = np.asarray(t.ints, dtype=np.float64)
        if k > 1:
            dims[i, :k] = v.shape
        ioff[i] = at
        flat = v.ravel(order="F")
        chunks.append(flat)
        at += flat.size
    tc_arr = np.array(tcs, dtype=np.int16).reshape(len(tcs), N, 2)
    return FlatHam(T, nbody, clusters, ops, parity, dims, delta, ioff,
                   np.concatenate(chunks) if chunks else np.zeros(0), len(tcs), tc_arr,
                   np.array(tc_start, dtype=np.int32), terms)


def cptr(a, ctype):
    import ctypes
    return a.ctypes.data_as(ctypes.POINTER(ctype))

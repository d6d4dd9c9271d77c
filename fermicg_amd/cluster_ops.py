"""Cluster-local many-body bases and transition-density tensors (ClusterOps).

Host-side input producer for the sigma build.  For every cluster this solves
the local FCI problem in each Fock sector (nα, nβ) and forms

    Γ_op[p(q r), I, J] = <I| o1_p o2_q o3_r |J>

for every operator string the ClusteredTerms use (reference
build_local_quantities.jl:40-191, tdm_* at :203-607).  Index order follows
the operator string, one index per operator, and each block is stored
column-major as (n^k, M_bra, M_ket) with the first operator index fastest
(build_local_quantities.jl:178-188); "H" and "Hcmf" are 2-D (M_bra, M_ket).

Local Fock-space convention: spin-orbital modes are ordered α_0..α_{n-1},
β_0..β_{n-1}; a determinant is the ordered product of its creation
operators on the vacuum, so a†_{pβ} picks up (-1)^(nα + #β below p).
Matrix elements of physical operators are independent of that choice up to
the phases of the eigenvectors, which is all the reference relies on.
Eigenvector phases are fixed so the largest-magnitude component is positive.
"""
from itertools import combinations

import numpy as np

from .integrals import InCoreInts, subset

# Operator strings carried in ClusterOps.  'A'/'a' create/annihilate α,
# 'B'/'b' create/annihilate β.  The id of an op is its index here; the C ABI
# uses the same numbering.
OP_NAMES = [
    "H", "Hcmf",
    "A", "a", "B", "b",
    "AA", "aa", "BB", "bb", "Aa", "Bb", "Ab", "Ba", "AB", "ba", "BA", "ab",
    "AAa", "Aaa", "BBb", "Bbb", "ABa", "Aba", "ABb", "Bba",
]
OP_ID = {s: i for i, s in enumerate(OP_NAMES)}
STRING_OPS = OP_NAMES[2:]


def op_delta(opstring):
    """Fock-sector change (Δnα, Δnβ) produced by an operator string."""
    da = opstring.count("A") - opstring.count("a")
    db = opstring.count("B") - opstring.count("b")
    return da, db


class LocalFockSpace:
    """Determinants of one cluster, grouped by sector, with dense matrices of
    single creation/annihilation operators between sectors."""

    def __init__(self, n):
        self.n = n
        self.dets = {}
        self.index = {}
        for na in range(n + 1):
            for nb in range(n + 1):
                masks = []
                for ob in combinations(range(n), nb):
                    for oa in combinations(range(n), na):
                        m = 0
                        for p in oa:
                            m |= 1 << p
                        for p in ob:
                            m |= 1 << (n + p)
                        masks.append(m)
                masks = np.array(sorted(masks), dtype=np.int64)
                self.dets[(na, nb)] = masks
                self.index[(na, nb)] = {int(m): i for i, m in enumerate(masks)}
        self._cache = {}

    def dim(self, sec):
        return len(self.dets[sec])

    def single(self, kind, p, sec_ket):
        """Dense matrix of one operator from `sec_ket`; returns (sec_bra, M)
        or None when the target sector is outside the Fock space."""
        key = (kind, p, sec_ket)
        if key in self._cache:
            return self._cache[key]
        n = self.n
        na, nb = sec_ket
        if kind == "A":
            sec_bra, mode, create = (na + 1, nb), p, True
        elif kind == "a":
            sec_bra, mode, create = (na - 1, nb), p, False
        elif kind == "B":
            sec_bra, mode, create = (na, nb + 1), n + p, True
        elif kind == "b":
            sec_bra, mode, create = (na, nb - 1), n + p, False
        else:
            raise ValueError(kind)
        if sec_bra not in self.dets:
            self._cache[key] = None
            return None
        M = np.zeros((self.dim(sec_bra), self.dim(sec_ket)))
        idx_bra = self.index[sec_bra]
        below = (1 << mode) - 1
        for j, m in enumerate(self.dets[sec_ket]):
            m = int(m)
            occ = (m >> mode) & 1
            if create == bool(occ):
                continue
            sign = -1.0 if bin(m & below).count("1") % 2 else 1.0
            m2 = m ^ (1 << mode)
            M[idx_bra[m2], j] = sign
        self._cache[key] = (sec_bra, M)
        return self._cache[key]

    def hamiltonian(self, h1, h2, sec):
        """Dense local Hamiltonian in sector `sec` (no constant)."""
        n = self.n
        d = self.dim(sec)
        E = np.zeros((n, n, d, d))
        for p in range(n):
            for q in range(n):
                for c, a in (("A", "a"), ("B", "b")):
                    r1 = self.single(a, q, sec)
                    if r1 is None:
                        continue
                    mid, Aq = r1
                    r2 = self.single(c, p, mid)
                    if r2 is None:
                        continue
                    E[p, q] += r2[1] @ Aq
        hp = h1 - 0.5 * np.einsum("prrq->pq", h2)
        H = np.einsum("pq,pqij->ij", hp, E)
        En = E.reshape(n * n, d, d)
        V = h2.reshape(n * n, n * n)
        tmp = np.einsum("xy,yjk->xjk", V, En)
        H += 0.5 * np.einsum("xij,xjk->ik", En, tmp)
        return 0.5 * (H + H.T)

    def apply_string(self, opstring, sec_ket, X):
        """Apply an operator string (all index tuples) to the columns of X.

        Returns (sec_bra, Y) with Y[p1,..,pk, :, :] = o1_p1 ... ok_pk X, or
        None when any intermediate sector does not exist."""
        n = self.n
        sec = sec_ket
        Y = X[None]  # (batch, dim, m)
        for ch in reversed(opstring):
            blocks = []
            sec_next = None
            for p in range(n):
                r = self.single(ch, p, sec)
                if r is None:
                    return None
                sec_next, M = r
                blocks.append(np.matmul(M, Y))  # M applied to every batch slice
            # new leading axis is this operator's index; earlier operators
            # (to the left) will be prepended, so stack as (p, batch, ...)
            Y = np.stack(blocks, axis=0).reshape(n * Y.shape[0], -1, X.shape[1])
            sec = sec_next
        return sec, Y.reshape((n,) * len(opstring) + Y.shape[1:])


def _fix_phase(vecs):
    for k in range(vecs.shape[1]):
        i = np.argmax(np.abs(vecs[:, k]))
        if vecs[i, k] < 0:
            vecs[:, k] *= -1.0
    return vecs


class ClusterBasis:
    """Lowest states of the local FCI problem per Fock sector
    (reference compute_cluster_eigenbasis, build_local_quantities.jl:1090-1180)."""

    def __init__(self, idx, orbs, space, vectors, energies):
        self.idx = idx
        self.orbs = list(orbs)
        self.space = space
        self.vectors = vectors      # sector -> (dim, M) array
        self.energies = energies    # sector -> (M,)

    def sectors(self):
        return list(self.vectors.keys())

    def __contains__(self, sec):
        return sec in self.vectors

    def __getitem__(self, sec):
        return self.vectors[sec]

    def n_states(self, sec):
        return self.vectors[sec].shape[1]


def compute_cluster_eigenbasis(ints: InCoreInts, clusters, max_roots=10, sectors=None,
                               rdm1a=None, rdm1b=None):
    """Per-cluster local bases.  `clusters` is a list of 0-based orbital
    lists; `sectors` optionally restricts each cluster's Fock sectors
    (list per cluster of (nα,nβ)); default is every sector of the cluster."""
    bases = []
    for ci, orbs in enumerate(clusters):
        n = len(orbs)
        space = LocalFockSpace(n)
        li = subset(ints, orbs, rdm1a, rdm1b)
        secs = sectors[ci] if sectors is not None else [
            (na, nb) for na in range(n + 1) for nb in range(n + 1)]
        vecs, ens = {}, {}
        for sec in secs:
            H = space.hamiltonian(li.h1, li.h2, sec)
            w, v = np.linalg.eigh(H)
            m = min(max_roots, len(w))
            vecs[sec] = _fix_phase(np.ascontiguousarray(v[:, :m]))
            ens[sec] = w[:m]
        bases.append(ClusterBasis(ci, orbs, space, vecs, ens))
    return bases


def compute_cluster_ops(bases, ints: InCoreInts):
    """ClusterOps for every cluster: list of dict op -> {(sec_bra, sec_ket): Γ}
    (reference compute_cluster_ops, build_local_quantities.jl:40-191).  The
    "H" blocks are V_bra' H V_ket in each sector (tdm_H, :203-228)."""
    all_ops = []
    for cb in bases:
        space = cb.space
        n = space.n
        li = subset(ints, cb.orbs)
        ops = {name: {} for name in OP_NAMES if name != "Hcmf"}
        for sec in cb.sectors():
            V = cb[sec]
            H = space.hamiltonian(li.h1, li.h2, sec)
            ops["H"][(sec, sec)] = np.asfortranarray(V.T @ H @ V)
        for name in STRING_OPS:
            k = len(name)
            da, db = op_delta(name)
            for sec_ket in cb.sectors():
                sec_bra = (sec_ket[0] + da, sec_ket[1] + db)
                if sec_bra not in cb:
                    continue
                Vb = cb[sec_bra]
                r = space.apply_string(name, sec_ket, cb[sec_ket])
                if r is None:
                    # an intermediate sector is empty: the block exists and is zero
                    G = np.zeros((n,) * k + (Vb.shape[1], cb.n_states(sec_ket)))
                else:
                    G = np.matmul(Vb.T, r[1])  # Vb' applied to every (p1..pk) slice
                # (p1..pk, I, J) -> column-major (n^k, Mb, Mk), p1 fastest
                Gf = np.asfortranarray(G).reshape((n ** k, Vb.shape[1], cb.n_states(sec_ket)), order="F")
                ops[name][(sec_bra, sec_ket)] = np.asfortranarray(Gf)
        all_ops.append(ops)
    return all_ops


def add_cmf_operators(all_ops, bases, ints: InCoreInts, rdm1a, rdm1b):
    """Embedded local Hamiltonians "Hcmf" (reference add_cmf_operators!,
    build_local_quantities.jl:616-651)."""
    for cb, ops in zip(bases, all_ops):
        li = subset(ints, cb.orbs, rdm1a, rdm1b)
        d = {}
        for sec in cb.sectors():
            V = cb[sec]
            H = cb.space.hamiltonian(li.h1, li.h2, sec) + li.h0 * np.eye(len(V))
            d[(sec, sec)] = np.asfortranarray(V.T @ H @ V)
        ops["Hcmf"] = d
    return all_ops


def product_state_rdm1(bases, n_orb, fspace):
    """Spin-resolved 1-RDMs of the product of each cluster's lowest state in
    sector `fspace[c]` (block diagonal; used as the CMF embedding density)."""
    Da = np.zeros((n_orb, n_orb)); Db = np.zeros((n_orb, n_orb))
    for cb, sec in zip(bases, fspace):
        n = cb.space.n
        v = cb[sec][:, 0]
        orbs = cb.orbs
        for p in range(n):
            for q in range(n):
                for c, a, D in (("A", "a", Da), ("B", "b", Db)):
                    r1 = cb.space.single(a, q, sec)
                    if r1 is None:
                        continue
                    mid, Aq = r1
                    r2 = cb.space.single(c, p, mid)
                    if r2 is None:
                        continue
                    D[orbs[p], orbs[q]] += v @ (r2[1] @ (Aq @ v))
    return Da, Db

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
    # spin and one-particle operators (build_local_quantities.jl:81-124,
    # 168-172): S+_p = a'_pα a_pβ and S-_p = a'_pβ a_pα per orbital, Sz and
    # the local S^2 (both on the sector's states), E1 = Aa + Bb
    "S+", "S-", "Sz", "S2", "E1",
]
OP_ID = {s: i for i, s in enumerate(OP_NAMES)}
STRING_OPS = OP_NAMES[2:26]


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


def _spin_ladder(space, V, sec, lower):
    """S- (lower) or S+ applied to the columns of V in sector `sec`: the
    images in the neighbouring sector, columns that vanish (states whose S
    is below the new |Ms|) dropped and the rest normalised; returns
    (vectors, sector) or (empty, sector) when the sector does not exist."""
    n = space.n
    na, nb = sec
    tgt = (na - 1, nb + 1) if lower else (na + 1, nb - 1)
    if not (0 <= tgt[0] <= n and 0 <= tgt[1] <= n):
        return np.zeros((0, 0)), tgt
    W = np.zeros((space.dim(tgt), V.shape[1]))
    for p in range(n):
        if lower:    # a'_pβ a_pα
            r1 = space.single("a", p, sec)
            r2 = space.single("B", p, r1[0]) if r1 is not None else None
        else:        # a'_pα a_pβ
            r1 = space.single("b", p, sec)
            r2 = space.single("A", p, r1[0]) if r1 is not None else None
        if r1 is None or r2 is None:
            continue
        W += r2[1] @ (r1[1] @ V)
    nrm = np.linalg.norm(W, axis=0)
    keep = nrm > 1e-10 * max(1.0, float(np.abs(V).max()) if V.size else 1.0)
    return W[:, keep] / nrm[keep], tgt


def compute_cluster_eigenbasis_spin(ints: InCoreInts, clusters, rdm1a, rdm1b, delta_elec,
                                    ref_fock, max_roots=10):
    """Spin-adapted local bases (compute_cluster_eigenbasis_spin,
    build_local_quantities.jl:899-1068).  For each cluster and each electron
    count within delta_elec[c] of its count in ref_fock, the local FCI of
    the embedded Hamiltonian is solved in the highest-Ms sector
    (n/2 + n%2, n/2), its lowest max_roots states kept; then S- is applied
    repeatedly (min(nα, n_orb - nβ) times) and S+ likewise
    (min(nβ, n_orb - nα) times), each image sector holding the normalised
    images of the states that survive (the spin multiplet partners, so every
    Ms component of a multiplet carries the same phase convention), with
    energies <v|H|v>."""
    bases = []
    for ci, orbs in enumerate(clusters):
        n = len(orbs)
        space = LocalFockSpace(n)
        li = subset(ints, orbs, rdm1a, rdm1b)
        ni = ref_fock[ci][0] + ref_fock[ci][1]
        d = delta_elec[ci]
        sectors = [(nj // 2 + nj % 2, nj // 2) for nj in range(ni - d, ni + d + 1)
                   if 0 <= nj <= 2 * n]
        vecs, ens = {}, {}
        for sec in sectors:
            H = space.hamiltonian(li.h1, li.h2, sec)
            w, v = np.linalg.eigh(H)
            m = min(max_roots, len(w))
            vecs[sec] = _fix_phase(np.ascontiguousarray(v[:, :m]))
            ens[sec] = w[:m]
            for lower, steps in ((True, min(sec[0], n - sec[1])), (False, min(sec[1], n - sec[0]))):
                vi, si = vecs[sec], sec
                for _ in range(steps):
                    vi, si = _spin_ladder(space, vi, si, lower)
                    if vi.shape[1] == 0:
                        break
                    Hi = space.hamiltonian(li.h1, li.h2, si)
                    vecs[si] = np.ascontiguousarray(vi)
                    ens[si] = np.einsum("ik,ij,jk->k", vi, Hi, vi)
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
        _add_spin_ops(ops, cb)
        all_ops.append(ops)
    return all_ops


def _add_spin_ops(ops, cb):
    """S+, S-, Sz, S2 and E1 of one cluster (build_local_quantities.jl:81-124,
    168-172): S+[p, I, J] = Ab[p, p, I, J], S-[p, I, J] = Ba[p, p, I, J],
    Sz = (nα - nβ)/2 on each sector's states (shape (1, M, M)), S2 = the
    sector's local S^2 = S- S+ + Sz (Sz + 1) between its states (2-D like
    H, tdm_S2 :233-257), E1 = Aa + Bb."""
    n = cb.space.n
    for name, src in (("S+", "Ab"), ("S-", "Ba")):
        d = {}
        for (sb, sk), G in ops[src].items():
            G4 = G.reshape(n, n, G.shape[1], G.shape[2], order="F")
            d[(sb, sk)] = np.asfortranarray(np.stack([G4[p, p] for p in range(n)], axis=0))
        ops[name] = d
    sz, s2, e1 = {}, {}, {}
    for sec in cb.sectors():
        V = cb[sec]
        M = V.shape[1]
        m = 0.5 * (sec[0] - sec[1])
        sz[(sec, sec)] = np.asfortranarray((m * np.eye(M)).reshape(1, M, M))
        S2 = (m * (m + 1.0)) * np.eye(V.shape[0])
        up = (sec[0] + 1, sec[1] - 1)
        if sec[1] > 0 and sec[0] < n:
            Sp = np.zeros((len(cb.space.dets[up]), V.shape[0]))
            for p in range(n):
                mid, Bq = cb.space.single("b", p, sec)
                _, Ap = cb.space.single("A", p, mid)
                Sp += Ap @ Bq
            S2 += Sp.T @ Sp  # S- = (S+)'
        s2[(sec, sec)] = np.asfortranarray(V.T @ S2 @ V)
        e1[(sec, sec)] = np.asfortranarray(ops["Aa"][(sec, sec)] + ops["Bb"][(sec, sec)])
    ops["Sz"], ops["S2"], ops["E1"] = sz, s2, e1


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

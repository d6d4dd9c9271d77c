"""ClusteredTerm / ClusteredOperator: the Hamiltonian split into 1-4-cluster
terms (reference type_ClusteredTerm.jl, type_ClusteredOperator.jl).

H = sum_pq h_pq p'q + 1/2 sum_pqrs (pr|qs) p'q's r   is regrouped by which
clusters the four (two) orbital indices live on.  For each assignment of
clusters to operator positions the positions are stably sorted by cluster
(bubble_sort, Utils.jl:47-64), the integrals are permuted with them and
multiplied by the permutation sign, and the operators of each cluster are
concatenated into that cluster's operator string.  A few strings are
rewritten into the canonical ones ClusterOps holds ("BA"->"AB", "ba"->"ab",
"BAa"->"ABa", "BAb"->"ABb", "Bab"->"Bba", "Aab"->"Aba"), swapping the two
integral axes and flipping the sign (type_ClusteredOperator.jl:275-316,
446-472).  Terms with equal (ops, clusters) are then merged by summing their
integrals (unique!, :724-806).  Terms are grouped by the TransferConfig (the
per-cluster (Δnα, Δnβ)) they cause.
"""
from dataclasses import dataclass, field
from itertools import permutations
from math import prod

import numpy as np

from .cluster_ops import OP_ID


def bubble_sort(seq):
    """Stable sort permutation and swap count (Utils.jl:47-64), 0-based."""
    b = list(seq)
    perm = list(range(len(b)))
    swaps = 0
    for j in range(len(b)):
        for i in range(len(b) - 1 - j):
            if b[i] > b[i + 1]:
                swaps += 1
                b[i], b[i + 1] = b[i + 1], b[i]
                perm[i], perm[i + 1] = perm[i + 1], perm[i]
    return perm, swaps


@dataclass
class ClusteredTerm:
    ops: tuple          # operator string per active cluster
    delta: tuple        # (Δnα, Δnβ) per active cluster
    parity: tuple       # len(op) % 2 per active cluster
    clusters: tuple     # active cluster indices, 0-based, ascending
    ints: np.ndarray    # column-major; ndim == nbody (1B: shape (1,))
    cache: dict = field(default_factory=dict, repr=False)

    @property
    def nbody(self):
        return len(self.clusters)

    def key(self):
        return "".join("%s(%d)" % (o, c) for o, c in zip(self.ops, self.clusters))


class ClusteredOperator:
    """OrderedDict TransferConfig -> [ClusteredTerm] (type_ClusteredOperator.jl:5-8)."""

    def __init__(self, n_clusters):
        self.n_clusters = n_clusters
        self.trans = {}

    def add(self, tc, term):
        self.trans.setdefault(tc, []).append(term)

    def __iter__(self):
        return iter(self.trans.items())

    def __len__(self):
        return len(self.trans)

    def __getitem__(self, tc):
        return self.trans[tc]

    def __contains__(self, tc):
        return tc in self.trans

    def terms(self):
        for tc, ts in self.trans.items():
            for t in ts:
                yield tc, t

    def n_terms(self):
        return sum(len(v) for v in self.trans.values())


def _zero_tc(n):
    return tuple((0, 0) for _ in range(n))


def _tc_replace(n, idx, focks):
    tc = [(0, 0)] * n
    for i, f in zip(idx, focks):
        tc[i] = (int(f[0]), int(f[1]))
    return tuple(tc)


_CANON_FIRST = {"BA": ("AB", (0, 1)), "ba": ("ab", (0, 1)), "BAa": ("ABa", (0, 1)),
                "BAb": ("ABb", (0, 1)), "Bab": ("Bba", (1, 2)), "Aab": ("Aba", (1, 2))}


def _canonicalize(opers, axes_of, v):
    """Rewrite non-canonical operator strings; `axes_of[c]` lists the integral
    axes (operator positions after sorting) that belong to cluster slot c.
    Swapping two adjacent operators of the same cluster flips the sign."""
    opers = list(opers)
    for c, s in enumerate(opers):
        if s in _CANON_FIRST:
            new, (i, j) = _CANON_FIRST[s]
            ax = axes_of[c]
            perm = list(range(v.ndim))
            perm[ax[i]], perm[ax[j]] = perm[ax[j]], perm[ax[i]]
            v = -np.transpose(v, perm)
            opers[c] = new
    return opers, v


def _emit(H, clus, sorted_cl, opers_sorted, focks_sorted, v, reshape):
    n = H.n_clusters
    slot = {c: k for k, c in enumerate(clus)}
    opers = [""] * len(clus)
    fock = [[0, 0] for _ in clus]
    axes_of = [[] for _ in clus]
    for pos, c in enumerate(sorted_cl):
        k = slot[c]
        opers[k] += opers_sorted[pos]
        fock[k][0] += focks_sorted[pos][0]
        fock[k][1] += focks_sorted[pos][1]
        axes_of[k].append(pos)
    opers, v = _canonicalize(opers, axes_of, v)
    if reshape:
        shape = [prod(v.shape[a] for a in axes_of[k]) for k in range(len(clus))]
        v = np.asfortranarray(v).reshape(shape, order="F")
    for o in opers:
        if o not in OP_ID:
            raise ValueError("operator string %s not in ClusterOps" % o)
    parity = tuple(len(o) % 2 for o in opers)
    term = ClusteredTerm(tuple(opers), tuple(tuple(f) for f in fock), parity,
                         tuple(clus), np.asfortranarray(v))
    H.add(_tc_replace(n, clus, fock), term)


_SPIN2 = [(["A", "A", "a", "a"], [(1, 0), (1, 0), (-1, 0), (-1, 0)]),
          (["B", "B", "b", "b"], [(0, 1), (0, 1), (0, -1), (0, -1)]),
          (["A", "B", "b", "a"], [(1, 0), (0, 1), (0, -1), (-1, 0)]),
          (["B", "A", "a", "b"], [(0, 1), (1, 0), (-1, 0), (0, -1)])]
_SPIN1 = [(["A", "a"], [(1, 0), (-1, 0)]), (["B", "b"], [(0, 1), (0, -1)])]


def _unique_perms(seq):
    seen, out = set(), []
    for p in permutations(seq):
        if p not in seen:
            seen.add(p)
            out.append(p)
    return out


def extract_ClusteredTerms(ints, clusters):
    """Build the ClusteredOperator for `ints` over `clusters` (0-based orbital
    lists).  Restates type_ClusteredOperator.jl:53-636."""
    n = len(clusters)
    orbs = [np.asarray(c, dtype=np.int64) for c in clusters]
    H = ClusteredOperator(n)
    zero = _zero_tc(n)
    H.trans[zero] = []
    for ci in range(n):
        H.add(zero, ClusteredTerm(("H",), ((0, 0),), (0,), (ci,), np.zeros(1)))

    def two_e(term_cl, clus):
        perm, swaps = bubble_sort(term_cl)
        sign = -1.0 if swaps % 2 else 1.0
        # positions (p, q, s, r) of p'q's r ; ints (pr|qs)
        vprqs = ints.h2[np.ix_(orbs[term_cl[0]], orbs[term_cl[3]], orbs[term_cl[1]], orbs[term_cl[2]])]
        vpqsr = np.transpose(vprqs, (0, 2, 3, 1))
        v = (0.5 * sign) * np.transpose(vpqsr, perm)
        sorted_cl = [term_cl[i] for i in perm]
        for ops4, focks4 in _SPIN2:
            _emit(H, clus, sorted_cl, [ops4[i] for i in perm], [focks4[i] for i in perm],
                  v.copy(), reshape=len(clus) < 4)

    # 2-body one-electron terms
    for ci in range(n):
        for cj in range(ci + 1, n):
            for term_cl in _unique_perms([ci, cj]):
                perm, swaps = bubble_sort(term_cl)
                sign = -1.0 if swaps % 2 else 1.0
                hpq = ints.h1[np.ix_(orbs[term_cl[0]], orbs[term_cl[1]])]
                h = sign * np.transpose(hpq, perm)
                sorted_cl = [term_cl[i] for i in perm]
                for ops2, focks2 in _SPIN1:
                    _emit(H, (ci, cj), sorted_cl, [ops2[i] for i in perm],
                          [focks2[i] for i in perm], h.copy(), reshape=False)
    # 2-body two-electron terms
    for ci in range(n):
        for cj in range(ci + 1, n):
            pats = (_unique_perms([ci, ci, cj, cj]) + _unique_perms([ci, cj, cj, cj])
                    + _unique_perms([ci, ci, ci, cj]))
            for term_cl in pats:
                two_e(term_cl, (ci, cj))
    # 3-body
    for ci in range(n):
        for cj in range(ci + 1, n):
            for ck in range(cj + 1, n):
                pats = (_unique_perms([ci, ci, cj, ck]) + _unique_perms([ci, cj, cj, ck])
                        + _unique_perms([ci, cj, ck, ck]))
                for term_cl in pats:
                    two_e(term_cl, (ci, cj, ck))
    # 4-body
    for ci in range(n):
        for cj in range(ci + 1, n):
            for ck in range(cj + 1, n):
                for cl in range(ck + 1, n):
                    for term_cl in _unique_perms([ci, cj, ck, cl]):
                        two_e(term_cl, (ci, cj, ck, cl))
    unique(H)
    return H


def unique(H):
    """Merge terms with equal (ops, clusters) by summing integrals."""
    for tc, terms in list(H.trans.items()):
        merged = {}
        for t in terms:
            k = t.key()
            if k in merged:
                merged[k].ints = merged[k].ints + t.ints
            else:
                merged[k] = ClusteredTerm(t.ops, t.delta, t.parity, t.clusters, t.ints.copy())
        H.trans[tc] = list(merged.values())
    return H


def extract_1body_operator(H, op_string="H"):
    """Only the 1-body terms, relabelled to `op_string` (e.g. "Hcmf");
    type_ClusteredOperator.jl:819-835."""
    out = ClusteredOperator(H.n_clusters)
    for tc, terms in H:
        for t in terms:
            if t.nbody == 1:
                out.add(tc, ClusteredTerm((op_string,), t.delta, t.parity, t.clusters, t.ints))
    return out


def extract_S2(n_clusters, cluster_sizes):
    """The total S^2 as a ClusteredOperator (extract_S2,
    type_ClusteredOperator.jl:644-715): per cluster the local "S2"
    (1-body), per cluster pair i < j the terms S+_i S-_j and S-_i S+_j
    (integrals of ones over the orbitals of both clusters, so each sums its
    per-orbital operators) and 2 Sz_i Sz_j, in the reference's order."""
    N = n_clusters
    out = ClusteredOperator(N)
    out.trans[_zero_tc(N)] = []
    for i in range(N):
        out.add(_zero_tc(N), ClusteredTerm(("S2",), ((0, 0),), (0,), (i,), np.ones(1)))
        for j in range(i + 1, N):
            ni, nj = cluster_sizes[i], cluster_sizes[j]
            for (oi, oj), (fi, fj) in ((("S+", "S-"), ((1, -1), (-1, 1))),
                                       (("S-", "S+"), ((-1, 1), (1, -1)))):
                out.add(_tc_replace(N, (i, j), (fi, fj)),
                        ClusteredTerm((oi, oj), (fi, fj), (0, 0), (i, j),
                                      np.ones((ni, nj), order="F")))
            out.add(_zero_tc(N), ClusteredTerm(("Sz", "Sz"), ((0, 0), (0, 0)), (0, 0), (i, j),
                                               2.0 * np.ones((1, 1), order="F")))
    return out

"""Determinant-space oracle (test infrastructure only; never on the product path).

An independent full-CI restatement used to pin the TPS sigma build without the
reference: Slater determinants of the whole system, alpha string x beta
string, with the textbook sigma

    sigma = sum_pq k_pq E_pq C + 1/2 sum_pqrs (pq|rs) E_pq E_rs C,
    k_pq  = h_pq - 1/2 sum_r (pr|rq)

(E_pq = sum_spin a'_p a_q).  Two uses:
  * FCI energies, checked against the reference fixtures' pinned e_fci
    (test/_testdata_cmf_h8.jld2, h9, hf_h6; generate_test_data_h12.jl:46-47);
  * the projection P' H P of the determinant Hamiltonian onto a set of
    tensor-product configurations, built from the local cluster vectors, which
    must equal the TPS Hamiltonian the sigma build applies
    (reference tps_ci_matvec, tpsci_outer.jl:462-527; build_full_H :10-50).

Global convention: modes alpha_0..alpha_{n-1}, beta_0..beta_{n-1}; a
determinant is the ordered product of its creation operators on the vacuum.
"""
from itertools import combinations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def strings(n, k):
    out = []
    for occ in combinations(range(n), k):
        m = 0
        for p in occ:
            m |= 1 << p
        out.append(m)
    return sorted(out)


def _popcount(x):
    return bin(x).count("1")


def excitation_matrices(n, k):
    """E[p][q] as sparse (n_str x n_str) matrices on strings of k electrons
    among n orbitals: E_pq |J> = sign |I>."""
    strs = strings(n, k)
    idx = {s: i for i, s in enumerate(strs)}
    ns = len(strs)
    E = [[None] * n for _ in range(n)]
    for p in range(n):
        for q in range(n):
            rows, cols, vals = [], [], []
            for j, s in enumerate(strs):
                if not (s >> q) & 1:
                    continue
                sign = -1 if _popcount(s & ((1 << q) - 1)) % 2 else 1
                s1 = s ^ (1 << q)
                if (s1 >> p) & 1:
                    continue
                sign *= -1 if _popcount(s1 & ((1 << p) - 1)) % 2 else 1
                s2 = s1 | (1 << p)
                rows.append(idx[s2]); cols.append(j); vals.append(float(sign))
            E[p][q] = sp.csr_matrix((vals, (rows, cols)), shape=(ns, ns))
    return strs, E


class DetFCI:
    def __init__(self, h1, h2, na, nb):
        self.n = h1.shape[0]
        self.na, self.nb = na, nb
        self.h1, self.h2 = h1, h2
        self.astr, self.Ea = excitation_matrices(self.n, na)
        self.bstr, self.Eb = excitation_matrices(self.n, nb)
        self.shape = (len(self.astr), len(self.bstr))
        self.k = h1 - 0.5 * np.einsum("prrq->pq", h2)

    @property
    def dim(self):
        return self.shape[0] * self.shape[1]

    def _E(self, p, q, C):
        # alpha acts on the row index, beta on the column index; a beta pair
        # passes the alpha string twice, so no extra sign.
        return self.Ea[p][q] @ C + (self.Eb[p][q] @ C.T).T

    def sigma(self, c):
        n = self.n
        C = c.reshape(self.shape)
        D = np.empty((n, n) + self.shape)
        for p in range(n):
            for q in range(n):
                D[p, q] = self._E(p, q, C)
        S = np.einsum("pq,pqij->ij", self.k, D)
        G = np.einsum("rspq,pqij->rsij", self.h2, D)
        for r in range(n):
            for s in range(n):
                S += 0.5 * self._E(r, s, G[r, s])
        return S.reshape(-1)

    def energies(self, nroots=1, tol=1e-12):
        d = self.dim
        if d <= 2500:
            H = np.column_stack([self.sigma(e) for e in np.eye(d)])
            return np.linalg.eigvalsh(0.5 * (H + H.T))[:nroots]
        op = spla.LinearOperator((d, d), matvec=self.sigma, dtype=float)
        w = spla.eigsh(op, k=nroots, which="SA", tol=tol, return_eigenvectors=False)
        return np.sort(w)

    def det_index(self, amask, bmask):
        ia = self.astr.index(amask)
        ib = self.bstr.index(bmask)
        return ia * self.shape[1] + ib


def _perm_parity(seq):
    seq = list(seq)
    par = 0
    for i in range(len(seq)):
        for j in range(i + 1, len(seq)):
            if seq[i] > seq[j]:
                par ^= 1
    return par


def tps_embedding(fci, clusters, local, fock_list, cfg_fock, cfg):
    """Columns = global determinant expansion of each TPS configuration.

    clusters : list of orbital lists (0-based)
    local    : per cluster, dict (na, nb) -> (masks, vectors) where masks are
               local determinant bitmasks (bit p = alpha orbital p of the
               cluster, bit n_c + p = beta orbital p) and vectors (n_det, M)
    fock_list: (n_fock, N, 2);  cfg_fock: (n_cfg,);  cfg: (n_cfg, N) 1-based
    The TPS state is the ordered product over clusters of each cluster's
    creation string (alpha then beta inside a cluster) on the vacuum.
    """
    N = len(clusters)
    n = fci.n
    Phi = np.zeros((fci.dim, len(cfg)))
    for col in range(len(cfg)):
        fock = fock_list[cfg_fock[col]]
        terms = [(1.0, [])]
        for c in range(N):
            masks, vecs = local[c][(int(fock[c][0]), int(fock[c][1]))]
            v = vecs[:, int(cfg[col][c]) - 1]
            new = []
            for w, parts in terms:
                for m, x in zip(masks, v):
                    if x != 0.0:
                        new.append((w * x, parts + [int(m)]))
            terms = new
        for w, parts in terms:
            modes = []
            amask = bmask = 0
            for c, m in enumerate(parts):
                nc = len(clusters[c])
                for p in range(nc):
                    if (m >> p) & 1:
                        modes.append(clusters[c][p])
                        amask |= 1 << clusters[c][p]
                for p in range(nc):
                    if (m >> (nc + p)) & 1:
                        modes.append(n + clusters[c][p])
                        bmask |= 1 << clusters[c][p]
            sign = -1.0 if _perm_parity(modes) else 1.0
            Phi[fci.det_index(amask, bmask), col] += sign * w
    return Phi


def projected_hamiltonian(fci, Phi):
    HPhi = np.column_stack([fci.sigma(Phi[:, j]) for j in range(Phi.shape[1])])
    return Phi.T @ HPhi

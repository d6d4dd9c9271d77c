"""In-core integrals and the seeded synthetic-integral recipe.

`InCoreInts` mirrors the container the reference takes from InCoreIntegrals
(``ints.h0``, ``ints.h1[p,q]``, ``ints.h2[p,q,r,s] = (pq|rs)`` in chemist
notation); every energy in this package is electronic, i.e. ``h0`` is never
added (the convention of every pinned value in the reference's test/).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class InCoreInts:
    h0: float
    h1: np.ndarray
    h2: np.ndarray

    @property
    def n_orb(self) -> int:
        return self.h1.shape[0]


def synthetic_ints(n_orb: int, n_aux: int, seed: int) -> InCoreInts:
    """Seeded integrals: h = diag(linspace(-2,0)) + symmetric N(0,0.1)
    off-diagonal; (pq|rs) = sum_L B[L,pq] B[L,rs] with B ~ N(0,0.3) made
    symmetric in (p,q), so the ERI tensor has 8-fold symmetry and is PSD.

    Draw order (numpy Generator(PCG64(seed))): first an (n,n) N(0,0.1) matrix
    whose strict upper triangle is mirrored, then the (L,n,n) B tensor.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    off = rng.normal(0.0, 0.1, size=(n_orb, n_orb))
    up = np.triu(off, 1)
    h1 = up + up.T + np.diag(np.linspace(-2.0, 0.0, n_orb))
    B = rng.normal(0.0, 0.3, size=(n_aux, n_orb, n_orb))
    B = 0.5 * (B + B.transpose(0, 2, 1))
    Bm = B.reshape(n_aux, n_orb * n_orb)
    h2 = (Bm.T @ Bm).reshape(n_orb, n_orb, n_orb, n_orb)
    return InCoreInts(0.0, np.ascontiguousarray(h1), np.ascontiguousarray(h2))


def subset(ints: InCoreInts, orbs, rdm1a=None, rdm1b=None) -> InCoreInts:
    """Integrals restricted to `orbs` (0-based).  With 1-RDMs, the orbitals
    outside `orbs` are frozen at that density and folded into h0/h1:

        h1'[p,q] = h1[p,q] + sum_{r,s env} (pq|rs) D[r,s] - (ps|rq) Dx[r,s]
        h0'      = sum_{rs env} h1[r,s] D[r,s]
                   + 1/2 sum (rs|tu) (D[r,s] D[t,u] - Da/Db same-spin exchange)

    with D = Da + Db and Dx = (Da + Db)/2 (spin-averaged exchange, exact for
    Da = Db).  This is the embedding used for the CMF local Hamiltonian.
    """
    orbs = np.asarray(orbs, dtype=np.int64)
    h1 = ints.h1[np.ix_(orbs, orbs)].copy()
    h2 = ints.h2[np.ix_(orbs, orbs, orbs, orbs)].copy()
    if rdm1a is None and rdm1b is None:
        return InCoreInts(0.0, h1, h2)
    n = ints.n_orb
    env = np.setdiff1d(np.arange(n), orbs)
    Da = np.zeros((n, n)); Db = np.zeros((n, n))
    Da[np.ix_(env, env)] = rdm1a[np.ix_(env, env)]
    Db[np.ix_(env, env)] = rdm1b[np.ix_(env, env)]
    D = Da + Db
    J = np.einsum("pqrs,rs->pq", ints.h2, D)
    Ka = np.einsum("psrq,rs->pq", ints.h2, Da)
    Kb = np.einsum("psrq,rs->pq", ints.h2, Db)
    h1 = h1 + J[np.ix_(orbs, orbs)] - 0.5 * (Ka + Kb)[np.ix_(orbs, orbs)]
    e_env = np.einsum("rs,rs->", ints.h1, D) + 0.5 * (
        np.einsum("rs,rs->", J, D) - np.einsum("rs,rs->", Ka, Da) - np.einsum("rs,rs->", Kb, Db))
    return InCoreInts(float(ints.h0 + e_env), h1, h2)


def load_fixture(path):
    """Load a decoded reference fixture (.npz written from the JLD2 files)."""
    d = np.load(path)
    sizes = d["cluster_sizes"]
    orbs = d["cluster_orbs"] - 1
    clusters, at = [], 0
    for s in sizes:
        clusters.append(list(orbs[at:at + s]))
        at += s
    out = {
        "ints": InCoreInts(float(d["h0"]), d["h1"], d["h2"]),
        "clusters": clusters,
        "init_fspace": [tuple(int(x) for x in r) for r in d["init_fspace"]],
    }
    for k in ("e_fci", "e_cmf", "d1a", "d1b"):
        if k in d:
            out[k] = d[k]
    return out

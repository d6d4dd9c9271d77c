"""Sparse-index algebra and the TPSCI vector container.

FockConfig   = tuple of (nα, nβ) per cluster     (type_FockConfig.jl:10-12)
ClusterConfig= tuple of 1-based state indices     (type_ClusterConfig.jl:7-10)
TransferConfig = FockConfig - FockConfig          (Indexing.jl:5-23)

`TPSCIstate` keeps the reference's flattening order (type_TPSCIstate.jl:
139-229): Fock blocks in insertion order, configurations in insertion order
inside each block, row = configuration, column = root.  It is stored flat:
`fock` (n_fock, N, 2) int16, `fock_start` (n_fock+1) int64, `cfg` (n_cfg, N)
int16 1-based, `coeffs` (n_cfg, R) float64.
"""
from itertools import product

import numpy as np


def fock_add(f, tc):
    return tuple((a[0] + b[0], a[1] + b[1]) for a, b in zip(f, tc))


def fock_sub(a, b):
    return tuple((x[0] - y[0], x[1] - y[1]) for x, y in zip(a, b))


def n_transferred(fock, ref):
    """Electrons gained by clusters relative to `ref` (sum of positive
    per-cluster, per-spin changes)."""
    return sum(max(0, f[0] - r[0]) + max(0, f[1] - r[1]) for f, r in zip(fock, ref))


def fock_configs_within(bases, ref_fock, max_transfer):
    """Fock configurations with the reference's total (Nα, Nβ), every sector
    present in the cluster bases, and at most `max_transfer` electrons moved
    between clusters.  Order: Iterators.product over each cluster's sectors
    (first cluster fastest), as expand_to_full_space! (tpsci_outer.jl:830-857).

    The transferred count splits into an alpha and a beta part, so the alpha
    and beta occupation tuples are enumerated separately (pruned by their
    totals and their own transfer) and then paired."""
    na = sum(f[0] for f in ref_fock)
    nb = sum(f[1] for f in ref_fock)
    N = len(bases)
    order = [cb.sectors() for cb in bases]
    secset = [set(o) for o in order]

    def spin_tuples(spin, total):
        vals = [sorted({s[spin] for s in order[c]}) for c in range(N)]
        lo = [0] * (N + 1); hi = [0] * (N + 1)
        for c in range(N):
            lo[c + 1] = lo[c] + vals[c][0]
            hi[c + 1] = hi[c] + vals[c][-1]
        out, cur = [], [0] * N

        def rec(c, rem, tr):
            if c < 0:
                if rem == 0:
                    out.append((tuple(cur), tr))
                return
            for v in vals[c]:
                t = tr + max(0, v - ref_fock[c][spin])
                r2 = rem - v
                if t > max_transfer or r2 < lo[c] or r2 > hi[c]:
                    continue
                cur[c] = v
                rec(c - 1, r2, t)

        rec(N - 1, total, 0)
        return out

    A = spin_tuples(0, na)
    B = spin_tuples(1, nb)
    out = []
    for ta, tra in A:
        for tb, trb in B:
            if tra + trb > max_transfer:
                continue
            f = tuple((ta[c], tb[c]) for c in range(N))
            if all(f[c] in secset[c] for c in range(N)):
                out.append(f)
    idx = [{s: i for i, s in enumerate(order[c])} for c in range(N)]
    out.sort(key=lambda f: tuple(idx[c][f[c]] for c in reversed(range(N))))
    return out


class TPSCIstate:
    def __init__(self, n_clusters, R=1):
        self.N = n_clusters
        self.R = R
        self.fock = np.zeros((0, n_clusters, 2), np.int16)
        self.fock_start = np.zeros(1, np.int64)
        self.cfg = np.zeros((0, n_clusters), np.int16)
        self.coeffs = np.zeros((0, R))

    @classmethod
    def from_blocks(cls, n_clusters, blocks, R=1):
        """`blocks` = list of (fock, [configs]) in insertion order."""
        s = cls(n_clusters, R)
        focks, starts, cfgs = [], [0], []
        for f, confs in blocks:
            focks.append(f)
            cfgs.extend(confs)
            starts.append(len(cfgs))
        s.fock = np.array(focks, np.int16).reshape(len(focks), n_clusters, 2)
        s.fock_start = np.array(starts, np.int64)
        s.cfg = np.array(cfgs, np.int16).reshape(len(cfgs), n_clusters)
        s.coeffs = np.zeros((len(cfgs), R))
        return s

    @classmethod
    def full_blocks(cls, bases, focks, R=1):
        """Every ClusterConfig of every Fock block (expand_each_fock_space!,
        tpsci_outer.jl:796-819: Iterators.product, first cluster fastest)."""
        N = len(bases)
        blocks = []
        for f in focks:
            dims = [bases[c].n_states(f[c]) for c in range(N)]
            confs = [tuple(reversed(t)) for t in product(*[range(1, d + 1) for d in reversed(dims)])]
            blocks.append((f, confs))
        return cls.from_blocks(N, blocks, R)

    def __len__(self):
        return self.cfg.shape[0]

    @property
    def n_fock(self):
        return self.fock.shape[0]

    def fock_of_cfg(self):
        return np.repeat(np.arange(self.n_fock), np.diff(self.fock_start))

    def get_vector(self):
        return self.coeffs.copy()

    def set_vector(self, v):
        v = np.asarray(v, dtype=np.float64)
        if v.ndim == 1:
            v = v[:, None]
        if v.shape[0] != len(self):
            raise ValueError("dimension mismatch")
        self.R = v.shape[1]
        self.coeffs = v.copy()

    def focks(self):
        return [tuple((int(a), int(b)) for a, b in f) for f in self.fock]

    def blocks(self):
        for fi in range(self.n_fock):
            yield fi, self.fock_start[fi], self.fock_start[fi + 1]

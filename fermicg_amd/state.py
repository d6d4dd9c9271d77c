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


# ---- keyed algebra on TPSCIstates (the reference's OrderedDict operations;
# a configuration is the key (Fock block, ClusterConfig), blocks and
# configurations keep insertion order) ----

def _state_blocks(s):
    """{fock: {config: coefficient row}} in insertion order."""
    out = {}
    focks = s.focks()
    for fi, a, b in s.blocks():
        d = out.setdefault(focks[fi], {})
        for i in range(a, b):
            d[tuple(int(x) for x in s.cfg[i])] = s.coeffs[i].copy()
    return out


def _from_dicts(N, R, blocks):
    s = TPSCIstate(N, R)
    focks, starts, cfgs, rows = [], [0], [], []
    for f, d in blocks.items():
        focks.append(f)
        for k, v in d.items():
            cfgs.append(k)
            rows.append(v)
        starts.append(len(cfgs))
    s.fock = np.array(focks, np.int16).reshape(len(focks), N, 2)
    s.fock_start = np.array(starts, np.int64)
    s.cfg = np.array(cfgs, np.int16).reshape(len(cfgs), N)
    s.coeffs = np.array(rows, np.float64).reshape(len(cfgs), R)
    return s


def copy_state(s):
    o = TPSCIstate(s.N, s.R)
    o.fock, o.fock_start, o.cfg, o.coeffs = s.fock.copy(), s.fock_start.copy(), s.cfg.copy(), s.coeffs.copy()
    return o


def add_states(s1, s2):
    """s1 + s2 by key (add!, src/type_TPSCIstate.jl:617-633): s2's new
    configurations follow s1's in their block, its new blocks follow."""
    b = _state_blocks(s1)
    for f, d in _state_blocks(s2).items():
        t = b.setdefault(f, {})
        for k, v in d.items():
            t[k] = t[k] + v if k in t else v
    return _from_dicts(s1.N, s1.R, b)


def scale_state(s, M):
    """Coefficients times an R x R' matrix (the reference's `v * S`)."""
    o = copy_state(s)
    o.coeffs = s.coeffs @ np.asarray(M, np.float64)
    o.R = o.coeffs.shape[1]
    return o


def zero_state(s):
    o = copy_state(s)
    o.coeffs[:] = 0.0
    return o


def clip_state(s, thresh=1e-5):
    """Drop configurations whose every coefficient is below thresh in
    magnitude, then empty blocks (clip!, src/type_TPSCIstate.jl:589-600)."""
    b = _state_blocks(s)
    out = {}
    for f, d in b.items():
        kept = {k: v for k, v in d.items() if not np.all(np.abs(v) < thresh)}
        if kept:
            out[f] = kept
    return _from_dicts(s.N, s.R, out)


def project_out_state(v, w):
    """Remove from v the configurations of w; a block of v that this empties
    is removed (project_out!, src/tpsci_outer.jl:869-889)."""
    b = _state_blocks(v)
    for f, d in _state_blocks(w).items():
        if f not in b:
            continue
        for k in d:
            b[f].pop(k, None)
        if not b[f]:
            del b[f]
    return _from_dicts(v.N, v.R, b)


def overlap_states(v1, v2):
    """R x R overlap over common configurations (overlap,
    src/type_TPSCIstate.jl:416-436)."""
    b1 = _state_blocks(v1)
    S = np.zeros((v1.R, v2.R))
    for f, d in _state_blocks(v2).items():
        if f not in b1:
            continue
        for k, c in d.items():
            if k in b1[f]:
                S += np.outer(b1[f][k], c)
    return S


def possible_spin_focksectors(ref_fock):
    """Fock configurations that spin-complete `ref_fock`
    (possible_spin_focksectors, type_FockConfig.jl:83-121): per cluster the
    sectors (nα, nβ), (nα-1, nβ+1), ... down to the spin-flipped one (or
    upwards when nα < nβ), combined by Iterators.product (first cluster
    fastest) and kept when the total Nα and Nβ equal the reference's."""
    spaces = []
    for na, nb in ref_fock:
        sp = [(na, nb)]
        a, b = na, nb
        for _ in range(abs(na - nb)):
            a, b = (a - 1, b + 1) if na > nb else (a + 1, b - 1)
            sp.append((a, b))
        spaces.append(sp)
    Na = sum(f[0] for f in ref_fock)
    Nb = sum(f[1] for f in ref_fock)
    out = []
    for t in product(*reversed(spaces)):
        f = tuple(reversed(t))
        if sum(x[0] for x in f) == Na and sum(x[1] for x in f) == Nb:
            out.append(f)
    return out


def add_spin_focksectors(state):
    """add_spin_focksectors (type_TPSCIstate.jl:660-671): for every Fock
    block of `state`, each spin-completing Fock block (the block itself
    included) is set to hold only the all-ground configuration with zero
    coefficients -- the reference re-creates each such block empty
    (add_fockconfig!, :92-95) before inserting that configuration, so a
    block that existed keeps its position and loses its other entries."""
    b = _state_blocks(state)
    gs = tuple([1] * state.N)
    for f in list(b.keys()):
        for g in possible_spin_focksectors(f):
            b[g] = {gs: np.zeros(state.R)}
    return _from_dicts(state.N, state.R, b)

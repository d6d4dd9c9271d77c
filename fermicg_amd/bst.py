"""Block-sparse Tucker (BST) states and their GPU sigma / Davidson.

BSTstate mirrors the reference's OrderedDict{FockConfig, OrderedDict{
TuckerConfig, Tucker}} (type_BSTstate.jl:13-18, hosvd.jl:15-19), flattened:
per block a Fock configuration, per cluster a TuckerConfig range (1-based,
inclusive) and a Tucker factor (range x rank, orthonormal columns, or the
identity), and R core tensors (column-major, first cluster fastest).
get_vector/set_vector order (type_BSTstate.jl:499-591): blocks in insertion
order, each core vec'd, one column per root.

Entry points (reference file:line):
    cache_hamiltonian(bra, ket, cluster_ops, clustered_ham)   tucker_inner.jl:90
    build_sigma(sigma, ci, cluster_ops, clustered_ham; cache) tucker_inner.jl:165
    ci_solve(ci, cluster_ops, clustered_ham; ...)             tucker_outer.jl:68
"""
from dataclasses import dataclass
from itertools import combinations

import numpy as np

from . import _lib
from .state import n_transferred


@dataclass
class FlatBST:
    N: int
    n_fock: int
    fock: np.ndarray      # int16 [n_fock, N, 2]
    n_box: int
    box_fock: np.ndarray  # int32 [n_box]
    lo: np.ndarray        # int16 [n_box, N]
    hi: np.ndarray        # int16 [n_box, N]
    rank: np.ndarray      # int16 [n_box, N]
    foff: np.ndarray      # int64 [n_box, N]  (-1 = identity)
    factors: np.ndarray   # float64 pool (column-major range x rank)
    var_off: np.ndarray = None  # int64 [n_box + 1] first variable of each box


# variables per box of the device engine: a larger Tucker block is cut into
# boxes along its slowest clusters' ranks (contiguous runs of its variables)
BOX_MAX = 64


def _split_block(rgs, facs):
    """Boxes of at most BOX_MAX variables covering one block in its variable
    order (first cluster fastest): (ranges, factors) per box.  The clusters
    before k keep the whole block's rank, cluster k is cut into runs of
    columns, every later cluster into single columns; an identity factor is
    cut by sub-ranges, an explicit factor by column subsets over the same
    range."""
    N = len(rgs)
    ranks = [(f.shape[1] if f is not None else hi - lo + 1) for (lo, hi), f in zip(rgs, facs)]
    if int(np.prod(ranks)) <= BOX_MAX:
        return [(list(rgs), list(facs))]
    k, pre = 0, 1
    while pre * ranks[k] <= BOX_MAX:
        pre *= ranks[k]
        k += 1
    chunk = BOX_MAX // pre

    def cut(c, a, n):
        (lo, hi), f = rgs[c], facs[c]
        if f is None:
            return (lo + a, lo + a + n - 1), None
        return (lo, hi), np.asfortranarray(f[:, a:a + n])

    out = []
    slow = [0] * N
    while True:
        for a in range(0, ranks[k], chunk):
            n = min(chunk, ranks[k] - a)
            br, bf = list(rgs[:k]), list(facs[:k])
            r, f = cut(k, a, n)
            br.append(r)
            bf.append(f)
            for c in range(k + 1, N):
                r, f = cut(c, slow[c], 1)
                br.append(r)
                bf.append(f)
            out.append((br, bf))
        c = k + 1
        while c < N:
            slow[c] += 1
            if slow[c] < ranks[c]:
                break
            slow[c] = 0
            c += 1
        if c >= N:
            return out


class BSTstate:
    def __init__(self, N, R=1):
        self.N = N
        self.R = R
        self.focks = []        # list of Fock tuples
        self.blocks = []       # list of (fock index, ranges[N] (lo,hi), factors[N] (array or None))
        self.coeffs = np.zeros((0, R))
        self._flat = None

    def add_block(self, fock, ranges, factors):
        if fock not in self._fidx:
            self._fidx[fock] = len(self.focks)
            self.focks.append(fock)
        self.blocks.append((self._fidx[fock], [tuple(r) for r in ranges], factors))
        self._flat = None

    @property
    def _fidx(self):
        if not hasattr(self, "_fidx_"):
            self._fidx_ = {f: i for i, f in enumerate(self.focks)}
        return self._fidx_

    def ranks(self, b):
        _, rgs, facs = self.blocks[b]
        return [(f.shape[1] if f is not None else hi - lo + 1) for (lo, hi), f in zip(rgs, facs)]

    def dims(self):
        return [int(np.prod(self.ranks(b))) for b in range(len(self.blocks))]

    def __len__(self):
        return int(sum(self.dims()))

    def var_offsets(self):
        return np.concatenate([[0], np.cumsum(self.dims())]).astype(np.int64)

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

    def zero(self):
        self.coeffs = np.zeros_like(self.coeffs)

    def flat(self):
        """The device layout: one box per block of at most BOX_MAX variables,
        larger blocks cut into several boxes (_split_block)."""
        if self._flat is not None:
            return self._flat
        N = self.N
        boxes = []
        for fi, rgs, facs in self.blocks:
            for br, bf in _split_block(rgs, facs):
                boxes.append((fi, br, bf))
        nb = len(boxes)
        lo = np.zeros((nb, N), np.int16)
        hi = np.zeros((nb, N), np.int16)
        rk = np.zeros((nb, N), np.int16)
        fo = -np.ones((nb, N), np.int64)
        var_off = np.zeros(nb + 1, np.int64)
        chunks, at = [], 0
        for b, (fi, rgs, facs) in enumerate(boxes):
            for c in range(N):
                lo[b, c], hi[b, c] = rgs[c]
                F = facs[c]
                if F is None:
                    rk[b, c] = rgs[c][1] - rgs[c][0] + 1
                else:
                    rk[b, c] = F.shape[1]
                    fo[b, c] = at
                    flat = np.asarray(F, np.float64).ravel(order="F")
                    chunks.append(flat)
                    at += flat.size
            var_off[b + 1] = var_off[b] + int(np.prod(rk[b].astype(np.int64)))
        self._flat = FlatBST(N, len(self.focks), np.array(self.focks, np.int16).reshape(-1, N, 2),
                             nb, np.array([b[0] for b in boxes], np.int32), lo, hi, rk, fo,
                             np.concatenate(chunks) if chunks else np.zeros(0), var_off)
        return self._flat

    def signature(self):
        f = self.flat()
        return hash((f.fock.tobytes(), f.box_fock.tobytes(), f.lo.tobytes(), f.hi.tobytes(),
                     f.rank.tobytes(), f.foff.tobytes(), f.factors.tobytes()))

    def copy_structure(self, R=None):
        o = BSTstate(self.N, self.R if R is None else R)
        o.focks = list(self.focks)
        o.blocks = list(self.blocks)
        o.coeffs = np.zeros((len(self), o.R))
        o._flat = self._flat
        return o


def orthonormal_factor(rng, n, r):
    q, _ = np.linalg.qr(rng.normal(size=(n, r)))
    return np.asfortranarray(q)


def bst_pq_state(bases, focks, ref_fock, max_q=2, max_rank=8, R=2, seed=0):
    """P/Q Tucker blocks per Fock configuration: P = 1:1 (rank 1, identity),
    Q = 2:M (random rank 1..min(max_rank,|Q|), orthonormal factor); blocks
    with at most `max_q` clusters in Q, enumerated all-P, then one-Q in
    cluster order, then pairs; cores N(0,1) * 10^-transferred."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = len(bases)
    st = BSTstate(N, R)
    cores = []
    for f in focks:
        M = [bases[c].n_states(f[c]) for c in range(N)]
        qc = [c for c in range(N) if M[c] >= 2]
        sets = [()]
        for k in range(1, max_q + 1):
            sets += list(combinations(qc, k))
        for s in sets:
            rgs, facs = [], []
            for c in range(N):
                if c in s:
                    nq = M[c] - 1
                    r = int(rng.integers(1, min(max_rank, nq) + 1))
                    rgs.append((2, M[c]))
                    facs.append(orthonormal_factor(rng, nq, r))
                else:
                    rgs.append((1, 1))
                    facs.append(None)
            st.add_block(f, rgs, facs)
            d = int(np.prod([F.shape[1] if F is not None else 1 for F in facs]))
            cores.append(rng.normal(size=(d, R)) * 10.0 ** (-n_transferred(f, ref_fock)))
    st.coeffs = np.concatenate(cores, axis=0)
    return st


def bst_full_identity(bases, focks, R=1):
    """One block per Fock configuration covering every state with identity
    factors (the BS limit; equals the TPSCI full-block space)."""
    N = len(bases)
    st = BSTstate(N, R)
    for f in focks:
        st.add_block(f, [(1, bases[c].n_states(f[c])) for c in range(N)], [None] * N)
    st.coeffs = np.zeros((len(st), R))
    return st


def embedding(st, bases, tps_state):
    """Matrix U mapping BST variables into the full-block TPS space
    `tps_state` (fermicg_amd.state.TPSCIstate over the same Fock configs):
    column = Kronecker product of the block's factors placed in its ranges."""
    N = st.N
    fidx = {f: i for i, f in enumerate(tps_state.focks())}
    off = st.var_offsets()
    U = np.zeros((len(tps_state), len(st)))
    for b, (fi, rgs, facs) in enumerate(st.blocks):
        f = st.focks[fi]
        t = fidx[f]
        base = tps_state.fock_start[t]
        M = [bases[c].n_states(f[c]) for c in range(N)]
        mats = []
        for c in range(N):
            lo, hi = rgs[c]
            E = np.zeros((M[c], (facs[c].shape[1] if facs[c] is not None else hi - lo + 1)))
            E[lo - 1:hi, :] = facs[c] if facs[c] is not None else np.eye(hi - lo + 1)
            mats.append(E)
        K = mats[0]
        for c in range(1, N):
            K = np.kron(mats[c], K)  # first cluster fastest
        U[base:base + K.shape[0], off[b]:off[b + 1]] = K
    return U


# ---- GPU entry points ----

def _engine(engine):
    from .tpsci import default_engine
    return engine or default_engine()


def cache_hamiltonian(bra, ket, cluster_ops, clustered_ham, nbody=4,
                      precisions=("fp64", "fp32", "bf16x2"), engine=None):
    """Build (and keep resident) every Hamiltonian block of the space.  The
    reference caches per-term dense ops on the ClusteredTerms; here the
    blocks of H in the Tucker basis are built on the GPU and cached on the
    engine, keyed by the space.  bra and ket must share the basis."""
    if bra is not ket and bra.signature() != ket.signature():
        raise ValueError("cache_hamiltonian: bra and ket must have the same blocks")
    eng = _engine(engine)
    key = ("bst", id(cluster_ops), id(clustered_ham), nbody, ket.signature(),
           tuple(sorted(set(precisions) | {"fp64"})))
    pl = eng._plans.get(key)
    if pl is None:
        sp = _lib.DeviceSpace.bst(eng.ctx, ket.flat())
        pl = _lib.Plan.create(eng.ctx, eng.ham(clustered_ham, nbody), eng.ops(cluster_ops), sp,
                              nbody=nbody, precisions=tuple(sorted(set(precisions) | {"fp64"})))
        eng._plans[key] = pl
    return pl


def build_sigma(sigma_vector, ci_vector, cluster_ops, clustered_ham, nbody=4, cache=False,
                precision="fp64", engine=None):
    """sigma_vector.coeffs += H ci_vector.coeffs (reference build_sigma!, which
    accumulates into sigma_vector with add!, tucker_inner.jl:235-242)."""
    pl = cache_hamiltonian(ci_vector, sigma_vector, cluster_ops, clustered_ham, nbody=nbody,
                           precisions=("fp64", precision), engine=engine)
    s = pl.sigma(ci_vector.coeffs, precision)
    sigma_vector.coeffs = sigma_vector.coeffs + s
    return sigma_vector


def compute_s2(ci_vector, cluster_ops, engine=None):
    """<r|S^2|r> per root of a BST state (the report of ci_solve,
    src/tucker_outer.jl:199-211: build_sigma! with extract_S2, then the
    overlap); the S^2 plan is built for the call and freed."""
    eng = _engine(engine)
    sp = _lib.DeviceSpace.bst(eng.ctx, ci_vector.flat())
    pl = _lib.Plan.create(eng.ctx, eng.ham(eng.s2_operator(cluster_ops), 2),
                          eng.ops(cluster_ops), sp, nbody=2, precisions=("fp64",))
    s = pl.sigma(ci_vector.coeffs)
    del pl, sp
    return np.einsum("ir,ir->r", ci_vector.coeffs, s)


def _mode_product(core, ranks, mats):
    """core (prod(ranks) x R, each column a column-major tensor of shape
    ranks) with mode c multiplied by mats[c]' (core' = core x_c mats[c]')."""
    R = core.shape[1]
    out = np.empty((int(np.prod([m.shape[1] for m in mats])), R))
    for r in range(R):
        T = core[:, r].reshape(ranks, order="F")
        for c, F in enumerate(mats):
            T = np.moveaxis(np.tensordot(T, F, axes=([c], [0])), -1, c)
        out[:, r] = T.ravel(order="F")
    return out


def form_1body_operator_diagonal(ci_vector, cluster_ops, H0="Hcmf", pseudo_canon=True):
    """Zeroth-order diagonal for the BST preconditioner
    (form_1body_operator_diagonal!, src/tucker_pt.jl:284-355).  Per block
    and cluster, the local operator H0 on the block's TuckerConfig range is
    projected on the Tucker factor, U' H0 U.  With pseudo_canon it is
    diagonalised and the factor and core rotated by its eigenvectors
    (transform_basis!, src/hosvd.jl:621-631: U <- U F, core <- core x F',
    so the state is unchanged), the energies being its eigenvalues;
    without, they are its diagonal.  Returns (state in the rotated basis,
    diagonal of sum_c H0_c over its variables)."""
    out = BSTstate(ci_vector.N, ci_vector.R)
    out.focks = list(ci_vector.focks)
    off = ci_vector.var_offsets()
    hd = np.zeros(len(ci_vector))
    cores = []
    for b, (fi, rgs, facs) in enumerate(ci_vector.blocks):
        f = ci_vector.focks[fi]
        newf, rots, ens = [], [], []
        for c in range(ci_vector.N):
            lo, hi = rgs[c]
            Hi = np.asarray(cluster_ops[c][H0][(f[c], f[c])])[lo - 1:hi, lo - 1:hi]
            U = facs[c] if facs[c] is not None else np.eye(hi - lo + 1)
            Hu = U.T @ Hi @ U
            if pseudo_canon and U.shape[1] > 1:
                w, F = np.linalg.eigh(0.5 * (Hu + Hu.T))
                newf.append(np.asfortranarray(U @ F))
                rots.append(F)
                ens.append(w)
            else:
                newf.append(facs[c])
                rots.append(np.eye(U.shape[1]))
                ens.append(np.diag(Hu).copy())
        ranks = [len(e) for e in ens]
        D = ens[0]
        for c in range(1, ci_vector.N):
            D = np.add.outer(ens[c], D)  # first cluster fastest
        hd[off[b]:off[b + 1]] = D.ravel()
        core = ci_vector.coeffs[off[b]:off[b + 1]]
        cores.append(_mode_product(core, ranks, rots) if pseudo_canon else core.copy())
        out.blocks.append((fi, list(rgs), newf if pseudo_canon else list(facs)))
    out.coeffs = np.concatenate(cores, axis=0) if cores else np.zeros((0, ci_vector.R))
    return out, hd


def ci_solve(ci_vector, cluster_ops, clustered_ham, conv_thresh=1e-5, max_ss_vecs=12,
             max_iter=40, lindep_thresh=1e-10, precond=True, nbody=4, precision="fp64",
             refresh_k=0, verbose=0, engine=None):
    """Davidson over the concatenated cores (reference ci_solve,
    src/tucker_outer.jl:68-215).  As there, with precond the Tucker factors
    are first rotated to the pseudo-canonical basis of the clusters' Hcmf
    and the Hcmf diagonal there preconditions the solve
    (form_1body_operator_diagonal!, :91-96); the cached Hamiltonian is built
    in that basis, and the eigenvectors are returned in it.  Returns
    (energies, BSTstate with the eigenvectors; .s2 holds <S^2> per root)."""
    vec = ci_vector
    hd = None
    if precond:
        if any("Hcmf" not in ops for ops in cluster_ops):
            raise ValueError("ci_solve(precond=True) needs the clusters' Hcmf (add_cmf_operators)")
        vec, hd = form_1body_operator_diagonal(ci_vector, cluster_ops, H0="Hcmf",
                                               pseudo_canon=True)
    pl = cache_hamiltonian(vec, vec, cluster_ops, clustered_ham, nbody=nbody,
                           precisions=("fp64", precision), engine=engine)
    e, v, stats = pl.davidson(vec.coeffs, conv_thresh=conv_thresh, max_iter=max_iter,
                              max_ss_vecs=max_ss_vecs, lindep_thresh=lindep_thresh,
                              precision=precision, refresh_k=refresh_k, precond=precond,
                              precond_diag=hd)
    if not stats["converged"]:
        # the reference prints its residuals and goes on; so does this, loudly
        import warnings
        warnings.warn("ci_solve: not converged after %d iterations (residuals %s, conv_thresh %g)"
                      % (stats["iterations"], np.array2string(np.asarray(stats["residual"]),
                                                              precision=3), conv_thresh),
                      RuntimeWarning, stacklevel=2)
    out = vec.copy_structure()
    out.set_vector(v)
    out.davidson_stats = stats
    out.s2 = compute_s2(out, cluster_ops, engine=engine)
    if verbose:
        print(" %5s %12s %12s" % ("Root", "Energy", "S2"))
        for r in range(len(e)):
            print(" %5d %12.8f %12.8f" % (r + 1, e[r], abs(out.s2[r])))
    return e, out

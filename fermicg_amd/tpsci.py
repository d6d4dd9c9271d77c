"""TPSCI entry points on the GPU, mirroring the reference's Julia API.

    tps_ci_matvec(ci_vector, cluster_ops, clustered_ham)      tpsci_outer.jl:462
    compute_diagonal(ci_vector, cluster_ops, clustered_ham)   tpsci_outer.jl:694
    tps_ci_davidson(ci_vector, cluster_ops, clustered_ham; …) tpsci_outer.jl:364
    tps_ci_direct(ci_vector, cluster_ops, clustered_ham; …)   tpsci_outer.jl:165
    compute_pt2_energy(ci_vector, cluster_ops, clustered_ham; …) tpsci_pt2_energy.jl:10
    open_matvec_thread(ci_vector, cluster_ops, clustered_ham; …) tpsci_matvec_thread.jl:11
    compute_pt1_wavefunction(ci_vector, cluster_ops, clustered_ham; …) tpsci_pt1_wavefunction.jl:11
    compute_qdpt_energy(ci_vector, cluster_ops, clustered_ham; …) tpsci_pt2_energy.jl:311
    tpsci_ci(ci_vector, cluster_ops, clustered_ham; …)        tpsci.jl:39

Same arguments, same meaning, same output layout (rows in the TPSCIstate's
flattening order, one column per root).  All arithmetic runs in
libfcg_hip.so; device-resident copies of ClusterOps, the ClusteredOperator
and the Hamiltonian blocks of each space are cached on an Engine and reused
across calls (the reference rebuilds nothing between Davidson iterations
either; its op cache is term.cache, tucker_inner.jl:90-159).
"""
import warnings

import numpy as np

from . import _lib
from .marshal import flatten_ham, flatten_ops


def cluster_sizes_of(cluster_ops):
    """Orbitals per cluster, read from the single-creation Γ blocks."""
    sizes = []
    for ops in cluster_ops:
        blk = next(iter(ops["A"].values()), None) if "A" in ops else None
        if blk is None:
            blk = next(iter(ops["B"].values()))
        sizes.append(blk.shape[0])
    return sizes


class Engine:
    """Device context plus caches of uploaded inputs and plans."""

    def __init__(self, device=0):
        self.ctx = _lib.Context(device)
        self._ops = {}
        self._ham = {}
        self._plans = {}
        self._local = {}

    def ops(self, cluster_ops):
        key = id(cluster_ops)
        hit = self._ops.get(key)
        if hit is None or hit[0] is not cluster_ops:
            fops = flatten_ops(cluster_ops, cluster_sizes_of(cluster_ops))
            hit = (cluster_ops, _lib.DeviceOps.upload(self.ctx, fops), fops)
            self._ops[key] = hit
        return hit[1]

    def ham(self, clustered_ham, nbody=4):
        key = (id(clustered_ham), nbody)
        hit = self._ham.get(key)
        if hit is None or hit[0] is not clustered_ham:
            fh = flatten_ham(clustered_ham, nbody)
            hit = (clustered_ham, _lib.DeviceHam.upload(self.ctx, fh, clustered_ham.n_clusters))
            self._ham[key] = hit
        return hit[1]

    def tpsci_plan(self, state, cluster_ops, clustered_ham, nbody=4, precisions=("fp64",),
                   matrix_free=False):
        """The space's plan: stored Hamiltonian blocks, or with matrix_free a
        plan that forms sigma per call (rank <= 1 spaces only)."""
        sig = (state.fock.tobytes(), state.fock_start.tobytes(), state.cfg.tobytes())
        key = (id(cluster_ops), id(clustered_ham), nbody, hash(sig), tuple(sorted(precisions)),
               bool(matrix_free))
        pl = self._plans.get(key)
        if pl is None:
            sp = _lib.DeviceSpace.tpsci(self.ctx, state)
            make = _lib.Plan.create_matrix_free if matrix_free else _lib.Plan.create
            pl = make(self.ctx, self.ham(clustered_ham, nbody), self.ops(cluster_ops), sp,
                      nbody=nbody, precisions=precisions)
            self._plans[key] = pl
        return pl

    def local_operator(self, clustered_ham, op_string):
        """The 1-body part of `clustered_ham` relabelled to `op_string`
        (extract_1body_operator), cached so its device plans are reused."""
        from .operator import extract_1body_operator
        key = (id(clustered_ham), op_string)
        hit = self._local.get(key)
        if hit is None or hit[0] is not clustered_ham:
            hit = (clustered_ham, extract_1body_operator(clustered_ham, op_string=op_string))
            self._local[key] = hit
        return hit[1]

    def s2_operator(self, cluster_ops):
        """extract_S2 for these clusters, cached (one object per cluster
        layout, so its device plans are reused)."""
        from .operator import extract_S2
        sizes = tuple(cluster_sizes_of(cluster_ops))
        key = ("S2", sizes)
        hit = self._local.get(key)
        if hit is None:
            hit = (None, extract_S2(len(sizes), sizes))
            self._local[key] = hit
        return hit[1]

    def space_diagonal(self, state, cluster_ops, clustered_ham):
        """H's diagonal over the configurations of `state` without a plan
        (fcg_space_diagonal)."""
        sp = _lib.DeviceSpace.tpsci(self.ctx, state)
        return _lib.space_diagonal(self.ctx, self.ham(clustered_ham), self.ops(cluster_ops), sp)

    def pair_plan(self, bra_state, ket_state, cluster_ops, clustered_ham, nbody=4):
        """Plan of H between two TPSCI spaces (rows: bra configurations,
        columns: ket configurations), not cached."""
        bra = _lib.DeviceSpace.tpsci(self.ctx, bra_state)
        ket = _lib.DeviceSpace.tpsci(self.ctx, ket_state)
        return _lib.Plan.create_pair(self.ctx, self.ham(clustered_ham, nbody),
                                     self.ops(cluster_ops), bra, ket, nbody=nbody)

    def bst_pair_sigma(self, bra, ket, cluster_ops, clustered_ham, nbody=4):
        """sigma over the blocks of BSTstate `bra` of H applied to BSTstate
        `ket` (n_bra x R), through a pair plan built for the call."""
        sb = _lib.DeviceSpace.bst(self.ctx, bra.flat())
        sk = _lib.DeviceSpace.bst(self.ctx, ket.flat())
        pl = _lib.Plan.create_pair(self.ctx, self.ham(clustered_ham, nbody), self.ops(cluster_ops),
                                   sb, sk, nbody=nbody)
        c = ket.coeffs
        out = np.zeros((len(bra), c.shape[1]))
        for r0 in range(0, c.shape[1], 16):
            r1 = min(c.shape[1], r0 + 16)
            out[:, r0:r1] = np.asarray(pl.sigma(c[:, r0:r1])).reshape(len(bra), r1 - r0)
        del pl
        return out

    def bst_ci_solve(self, vec, cluster_ops, clustered_ham, **kw):
        """bst.ci_solve on this engine; the plan it caches for the space is
        freed afterwards (the outer loop's spaces change every call)."""
        from .bst import ci_solve
        before = set(self._plans)
        try:
            return ci_solve(vec, cluster_ops, clustered_ham, engine=self, **kw)
        finally:
            for k in [k for k in self._plans if k not in before]:
                del self._plans[k]

    def open_sigma(self, ket_state, target_state, cluster_ops, clustered_ham, nbody=4,
                   thresh=0.0):
        """Screened open matvec of `ket_state` onto every configuration of
        `target_state` (fcg_open_sigma): (sigma [n_target x R], hit [n_target]
        bool, the configurations some term's element reached)."""
        bra = _lib.DeviceSpace.tpsci(self.ctx, target_state)
        ket = _lib.DeviceSpace.tpsci(self.ctx, ket_state)
        return _lib.open_sigma(self.ctx, self.ham(clustered_ham, nbody), self.ops(cluster_ops),
                               bra, ket, ket_state.coeffs, nbody=nbody, thresh=thresh)

    def open_sigma_sparse(self, ket_state, cluster_ops, clustered_ham, nbody=4, thresh=0.0):
        """Screened open matvec generated per (ket configuration, term)
        (fcg_open_sigma_sparse): (fock [n, N, 2], cfg [n, N], sigma [n, R])
        of the configurations some term's element reached."""
        ket = _lib.DeviceSpace.tpsci(self.ctx, ket_state)
        return _lib.open_sigma_sparse(self.ctx, self.ham(clustered_ham, nbody),
                                      self.ops(cluster_ops), ket, ket_state.coeffs,
                                      ket_state.N, nbody=nbody, thresh=thresh)

    def pt2(self, ci_vector, cluster_ops, clustered_ham, E0, nbody=4, H0="Hcmf", thresh=0.0,
            method="dense", precision="fp64"):
        """e2 of an orthonormal TPSCI state over its first-order space:
        method "dense" (fcg_tpsci_pt2, whole target Fock blocks) or "sparse"
        (fcg_tpsci_pt2_sparse, the configurations screened contributions
        reach).  Returns (e2 [R], statistics dict)."""
        from .cluster_ops import OP_NAMES
        space = _lib.DeviceSpace.tpsci(self.ctx, ci_vector)
        args = (self.ctx, self.ham(clustered_ham, nbody), self.ops(cluster_ops), space,
                ci_vector.coeffs, E0)
        kw = dict(nbody=nbody, h0_op=OP_NAMES.index(H0), thresh=thresh)
        if method == "sparse":
            return _lib.tpsci_pt2_sparse(*args, precision=precision, **kw)
        return _lib.tpsci_pt2(*args, **kw)

    def drop_plans(self, keep=()):
        """Free the cached plans (their device blocks) except those whose
        keys are in `keep`, keeping the uploaded operators: for loops whose
        spaces change every iteration."""
        for k in [k for k in self._plans if k not in keep]:
            del self._plans[k]

    def clear(self):
        self._local.clear()
        self._plans.clear()
        self._ham.clear()
        self._ops.clear()


_ENGINE = None


def default_engine():
    global _ENGINE
    if _ENGINE is None:
        _ENGINE = Engine(0)
    return _ENGINE


def tps_ci_matvec(ci_vector, cluster_ops, clustered_ham, precision="fp64", engine=None):
    """sigma = H c over the configurations of `ci_vector` (n_cfg x R)."""
    eng = engine or default_engine()
    pl = eng.tpsci_plan(ci_vector, cluster_ops, clustered_ham,
                        precisions=tuple(sorted({"fp64", precision})))
    return pl.sigma(ci_vector.coeffs, precision)


def compute_diagonal(ci_vector, cluster_ops, clustered_ham, engine=None):
    """Diagonal of `clustered_ham` over the configurations of `ci_vector`
    (compute_diagonal, src/tpsci_outer.jl:694-715): the zero-TransferConfig
    terms at bra = ket, one GPU thread per configuration, no plan built."""
    eng = engine or default_engine()
    return eng.space_diagonal(ci_vector, cluster_ops, clustered_ham)


def compute_s2(ci_vector, cluster_ops, engine=None):
    """<r|S^2|r> per root with the S^2 of extract_S2 (the report of
    tps_ci_direct / tps_ci_davidson, src/tpsci_outer.jl:287-294,436-442)."""
    eng = engine or default_engine()
    return compute_expectation_value(ci_vector, cluster_ops, eng.s2_operator(cluster_ops),
                                     engine=eng)


def _report_s2(out, e, cluster_ops, eng, verbose):
    s2 = compute_s2(out, cluster_ops, engine=eng)
    out.s2 = s2
    if verbose:
        print(" %5s %12s %12s" % ("Root", "Energy", "S2"))
        for r in range(len(e)):
            print(" %5d %12.8f %12.8f" % (r + 1, e[r], abs(s2[r])))


def tps_ci_davidson(ci_vector, cluster_ops, clustered_ham, conv_thresh=1e-5, lindep_thresh=1e-12,
                    max_ss_vecs=12, max_iter=40, precond=True, precision="fp64", refresh_k=0,
                    verbose=0, engine=None):
    """Lowest R eigenpairs in the space of `ci_vector`, starting from its
    coefficients.  Returns (energies, TPSCIstate with the eigenvectors);
    the state carries <S^2> per root as `.s2` (printed with verbose, as the
    reference prints it)."""
    eng = engine or default_engine()
    pl = eng.tpsci_plan(ci_vector, cluster_ops, clustered_ham,
                        precisions=tuple(sorted({"fp64", precision})))
    e, v, stats = pl.davidson(ci_vector.coeffs, conv_thresh=conv_thresh, max_iter=max_iter,
                              max_ss_vecs=max_ss_vecs, lindep_thresh=lindep_thresh,
                              precision=precision, refresh_k=refresh_k, precond=precond)
    if not stats["converged"]:
        # the reference prints its residuals and returns; so does this, loudly
        warnings.warn("tps_ci_davidson: not converged after %d iterations (residuals %s, "
                      "conv_thresh %g)" % (stats["iterations"],
                                           np.array2string(np.asarray(stats["residual"]),
                                                           precision=3), conv_thresh),
                      RuntimeWarning, stacklevel=2)
    out = _copy_state(ci_vector)
    out.set_vector(v)
    out.davidson_stats = stats
    _report_s2(out, e, cluster_ops, eng, verbose)
    return e, out


def build_full_H(ci_vector, cluster_ops, clustered_ham, engine=None):
    """Dense H over the configurations of `ci_vector` (n x n, the order of
    its flattening), as build_full_H_parallel (src/tpsci_outer.jl:59-149)
    returns it: the GPU plan's FP64 sigma applied to columns of the identity,
    up to 16 at a time."""
    eng = engine or default_engine()
    pl = eng.tpsci_plan(ci_vector, cluster_ops, clustered_ham)
    n = len(ci_vector)
    H = np.zeros((n, n))
    for j0 in range(0, n, 16):
        j1 = min(n, j0 + 16)
        E = np.zeros((n, j1 - j0))
        E[np.arange(j0, j1), np.arange(j1 - j0)] = 1.0
        H[:, j0:j1] = pl.sigma(E)
    return H


def tps_ci_direct(ci_vector, cluster_ops, clustered_ham, conv_thresh=1e-5, lindep_thresh=1e-12,
                  max_ss_vecs=12, max_iter=40, shift=None, precond=True, H_old=None, v_old=None,
                  verbose=0, solver="davidson", engine=None):
    """Lowest R eigenpairs through the dense H of the space
    (tps_ci_direct, src/tpsci_outer.jl:165-305).  Returns (e0, TPSCIstate
    with the eigenvectors, H).

    H is build_full_H (GPU).  As in the reference, a space of at most 500
    configurations is diagonalised whole (the R lowest eigenpairs of H);
    a larger one iteratively from the state's coefficients: the GPU Davidson
    on the same plan, with conv_thresh, lindep_thresh, max_ss_vecs, max_iter
    and precond as given (the reference's default solver="davidson"; the
    other solvers and `shift` are accepted for call compatibility).  H_old and
    v_old, the reference's incremental rebuild, are accepted and not needed:
    the whole H is rebuilt from the cached plan."""
    eng = engine or default_engine()
    H = build_full_H(ci_vector, cluster_ops, clustered_ham, engine=eng)
    R = ci_vector.R
    if len(ci_vector) > 500:
        e0, out = tps_ci_davidson(ci_vector, cluster_ops, clustered_ham, conv_thresh=conv_thresh,
                                  lindep_thresh=lindep_thresh, max_ss_vecs=max_ss_vecs,
                                  max_iter=max_iter, precond=precond, verbose=verbose,
                                  engine=eng)
        return np.asarray(e0), out, H
    w, v = np.linalg.eigh(H)
    out = _copy_state(ci_vector)
    out.set_vector(v[:, :R])
    _report_s2(out, w[:R], cluster_ops, eng, verbose)
    return w[:R].copy(), out, H


def orthonormalize(v):
    """Modified Gram-Schmidt over the root columns, in place order
    (orthonormalize!, src/type_TPSCIstate.jl:569-582)."""
    v = np.array(v, dtype=np.float64, copy=True)
    v[:, 0] /= np.linalg.norm(v[:, 0])
    for r in range(1, v.shape[1]):
        for r0 in range(r):
            v[:, r] -= v[:, r0] * (v[:, r0] @ v[:, r])
        v[:, r] /= np.linalg.norm(v[:, r])
    return v


def compute_expectation_value(ci_vector, cluster_ops, clustered_ham, engine=None):
    """<r|H|r> per root (compute_expectation_value_parallel,
    src/tpsci_outer.jl:601-690): sigma on the GPU, one dot product per root."""
    eng = engine or default_engine()
    c = ci_vector.coeffs
    s = eng.tpsci_plan(ci_vector, cluster_ops, clustered_ham).sigma(c)
    return np.einsum("ir,ir->r", c, s)


def compute_pt2_energy(ci_vector, cluster_ops, clustered_ham, nbody=4, H0="Hcmf", E0=None,
                       thresh_foi=1e-9, prescreen=True, verbose=1, method="auto", engine=None,
                       precision="fp64"):
    """Second-order energy correction e2[r] of a TPSCI state over its
    first-order space (compute_pt2_energy, src/tpsci_pt2_energy.jl:10-175).

    As in the reference: the state is orthonormalised (Gram-Schmidt over
    roots), E0 = <r|H0|r> with H0 the clusters' local `H0` operators unless
    given, and e2[r] = sum_x sigma[x,r]^2 / (E0[r] - H0_diag[x]) over every
    configuration of the Fock blocks F_ket + TransferConfig that is not in
    the variational space.  sigma over those blocks runs on the GPU with the
    reference's screening at `thresh_foi` (a term's contribution from ket
    configuration j enters only when its magnitude exceeds
    thresh_foi / max_r |c[j, r]|); `prescreen` is accepted for call
    compatibility: the reference's bounds only skip contributions that fail
    that test.  `method`: "dense" forms sigma on every configuration of each
    target Fock block (few clusters), "sparse" only on the configurations
    screened contributions reach, generated per (ket configuration, term)
    as the reference does (many clusters); "auto" takes "sparse" above 4
    clusters.  Both give the same e2.  `precision` "fp32" runs the sparse
    engine's contraction in FP32 (every sum after it in FP64).  Returns e2
    (length R)."""
    eng = engine or default_engine()
    if method == "auto":
        method = "sparse" if ci_vector.N > 4 else "dense"
    ci = _copy_state(ci_vector)
    ci.set_vector(orthonormalize(ci.coeffs))
    if E0 is None:
        E0 = compute_expectation_value(ci, cluster_ops, eng.local_operator(clustered_ham, H0),
                                       engine=eng)
    E0 = np.asarray(E0, dtype=np.float64).reshape(-1)
    if precision != "fp64" and method != "sparse":
        raise ValueError("a reduced-precision PT2 contraction needs method='sparse'")
    extra = {} if precision == "fp64" else {"precision": precision}
    e2, stats = eng.pt2(ci, cluster_ops, clustered_ham, E0, nbody=nbody, H0=H0,
                        thresh=thresh_foi, method=method, **extra)
    if verbose:
        evar = compute_expectation_value(ci, cluster_ops, clustered_ham, engine=eng)
        if method == "sparse":
            print(" first-order space: %d configurations reached by %d contributions; %.3f s"
                  % (stats["n_first_order"], stats["n_contributions"], stats["seconds"]))
        else:
            print(" first-order space: %d Fock blocks, %d configurations; %.3f s"
                  % (stats["n_target_fock"], stats["n_first_order"], stats["seconds"]))
        print(" %5s %12s %12s" % ("Root", "E(0)", "E(2)"))
        for r in range(len(e2)):
            print(" %5d %12.8f %12.8f" % (r + 1, evar[r], evar[r] + e2[r]))
    compute_pt2_energy.last_stats = stats
    return e2


def first_order_focks(ci_vector, cluster_ops, clustered_ham):
    """Target Fock blocks of the open matvec, F_ket + TransferConfig for every
    Fock block of the state and every TransferConfig of the operator, kept
    when each cluster's electron counts lie in [0, n_orbitals] and its sector
    has an "H" block (open_matvec_thread job list, src/tpsci_matvec_thread.jl:
    25-52), in first-appearance order (ket blocks outer, TransferConfigs
    inner).  Formed with array operations over ket blocks x TransferConfigs,
    so that spaces of many clusters (thousands of each) take seconds."""
    sizes = np.array(cluster_sizes_of(cluster_ops), np.int64)
    N = len(cluster_ops)
    F = np.asarray(ci_vector.fock, np.int64).reshape(-1, N, 2)
    T = np.array([[tuple(x) for x in tc] for tc in clustered_ham.trans],
                 np.int64).reshape(-1, N, 2)
    if len(F) == 0 or len(T) == 0:
        return []
    top = int(sizes.max()) + 1
    has_h = np.zeros((N, top, top), bool)
    for c in range(N):
        for fa, fb in cluster_ops[c]["H"]:
            if fa == fb and 0 <= fa[0] < top and 0 <= fa[1] < top:
                has_h[c, fa[0], fa[1]] = True
    cidx = np.arange(N)
    # each candidate packed into one integer key (sector code a * top + b per
    # cluster) when N clusters fit 63 bits, else compared as byte rows
    packed = (top * top) ** N < (1 << 63)
    weights = np.array([(top * top) ** c for c in range(N)], np.int64) if packed else None
    seen = np.zeros(0, np.int64) if packed else None
    seen_rows = set()
    keep = []
    step = max(1, (1 << 22) // (len(T) * N))
    for i0 in range(0, len(F), step):
        cand = F[i0:i0 + step, None] + T[None]                      # (f, t, N, 2)
        inr = ((cand >= 0) & (cand <= sizes[None, None, :, None])).all(axis=3)
        a = np.clip(cand[..., 0], 0, top - 1)
        b = np.clip(cand[..., 1], 0, top - 1)
        ok = (inr & has_h[cidx[None, None, :], a, b]).all(axis=2)   # (f, t)
        rows = cand[ok].reshape(-1, 2 * N)
        if len(rows) == 0:
            continue
        if packed:
            key = ((rows[:, 0::2] * top + rows[:, 1::2]) * weights).sum(axis=1)
            _, first = np.unique(key, return_index=True)
            first = np.sort(first)
            new = first[~np.isin(key[first], seen)]
            seen = np.union1d(seen, key[new])
            keep.append(rows[new])
        else:
            for r in rows:
                t = r.tobytes()
                if t not in seen_rows:
                    seen_rows.add(t)
                    keep.append(r[None])
    if not keep:
        return []
    keep = np.concatenate(keep)
    return [tuple((int(r[2 * c]), int(r[2 * c + 1])) for c in range(N)) for r in keep]


def _full_space(cluster_ops, focks, R):
    """Every configuration of each Fock block, first cluster fastest
    (expand_each_fock_space!, tpsci_outer.jl:796-819)."""
    from .state import TPSCIstate
    N = len(cluster_ops)
    s = TPSCIstate(N, R)
    cfgs, starts = [], [0]
    for f in focks:
        dims = [cluster_ops[c]["H"][(f[c], f[c])].shape[0] for c in range(N)]
        grids = np.meshgrid(*[np.arange(1, d + 1, dtype=np.int16) for d in dims], indexing="ij")
        cfgs.append(np.stack([g.ravel(order="F") for g in grids], axis=1))
        starts.append(starts[-1] + cfgs[-1].shape[0])
    s.fock = np.array(focks, np.int16).reshape(len(focks), N, 2)
    s.fock_start = np.array(starts, np.int64)
    s.cfg = np.concatenate(cfgs).astype(np.int16) if cfgs else np.zeros((0, N), np.int16)
    s.coeffs = np.zeros((s.cfg.shape[0], R))
    return s


def open_matvec_thread(ci_vector, cluster_ops, clustered_ham, thresh=1e-9, prescreen=True,
                       nbody=4, engine=None, method="dense"):
    """sigma = H c on the first-order space of a TPSCI state
    (open_matvec_thread, src/tpsci_matvec_thread.jl:11-129), on the GPU.

    The target Fock blocks are first_order_focks; sigma over every
    configuration of them is one screened open matvec (fcg_open_sigma: a
    term's element from ket configuration j enters only when its magnitude
    exceeds thresh / max_r |c[j, r]|, the test of
    _collect_significant_thread!, :830-895, built and applied in one pass,
    nothing stored).  Returns a TPSCIstate holding every target Fock block
    (the reference adds each job's block even when nothing lands in it) and
    in it the configurations some term's element reached -- the keys the
    reference inserts, kept even where the contributions cancel -- with
    sigma as coefficients.  Configurations follow each block's enumeration
    order (first cluster fastest) and blocks their first appearance; the
    reference's order is its Dicts', so compare by key.  `prescreen` is
    accepted for call compatibility: the reference's bounds only skip
    elements that fail the same test.

    method "sparse" generates the same configurations and sigma per (ket
    configuration, term) with the reference's bounds
    (fcg_open_sigma_sparse), for first-order spaces whose target Fock
    blocks (or whose ket blocks x TransferConfigs job list) are too large to
    list: its TPSCIstate holds only the Fock blocks something reached, in
    ascending order of their sector codes, each block's configurations in
    enumeration order (compare by key, as with the reference's Dicts)."""
    from .state import TPSCIstate
    eng = engine or default_engine()
    R = ci_vector.R
    if method == "sparse":
        return _open_sparse_state(eng, ci_vector, cluster_ops, clustered_ham, nbody, thresh)
    if method != "dense":
        raise ValueError("open_matvec_thread: method is 'dense' or 'sparse'")
    focks = first_order_focks(ci_vector, cluster_ops, clustered_ham)
    target = _full_space(cluster_ops, focks, R)
    sig, keep = eng.open_sigma(ci_vector, target, cluster_ops, clustered_ham, nbody=nbody,
                               thresh=thresh)
    out = TPSCIstate(ci_vector.N, R)
    out.fock = target.fock.copy()
    kept_before = np.concatenate([[0], np.cumsum(keep)]).astype(np.int64)
    out.fock_start = kept_before[target.fock_start]
    out.cfg = target.cfg[keep].copy()
    out.coeffs = sig[keep].copy()
    return out


def _open_sparse_state(eng, ci_vector, cluster_ops, clustered_ham, nbody, thresh):
    """The sparse open matvec's configurations as a TPSCIstate: the Fock
    blocks reached, in ascending order of (alpha, beta) counts cluster by
    cluster, each block's configurations in enumeration order (first cluster
    fastest)."""
    from .state import TPSCIstate
    N, R = ci_vector.N, ci_vector.R
    fock, cfg, sig = eng.open_sigma_sparse(ci_vector, cluster_ops, clustered_ham, nbody=nbody,
                                           thresh=thresh)
    out = TPSCIstate(N, R)
    if len(cfg) == 0:
        out.coeffs = np.zeros((0, R))
        return out
    fk, blk = np.unique(fock.reshape(len(cfg), 2 * N), axis=0, return_inverse=True)
    blk = np.asarray(blk).reshape(-1)
    order = np.lexsort(tuple(cfg[:, c] for c in range(N)) + (blk,))
    out.fock = fk.reshape(-1, N, 2).astype(np.int16)
    out.fock_start = np.concatenate([[0], np.cumsum(np.bincount(blk, minlength=len(fk)))]
                                    ).astype(np.int64)
    out.cfg = cfg[order].astype(np.int16)
    out.coeffs = sig[order].copy()
    return out


def compute_pt1_wavefunction(ci_vector, cluster_ops, clustered_ham, nbody=4, H0="Hcmf", E0=None,
                             thresh_foi=1e-8, prescreen=False, verbose=1, threaded=True,
                             engine=None, open_method="dense"):
    """First-order wavefunction and e2 (compute_pt1_wavefunction,
    src/tpsci_pt1_wavefunction.jl:11-95): sigma = open_matvec_thread on the
    GPU, the state's own configurations projected out, and per root
    v_pt = sigma / (E0 / |c|^2 - Hd) with Hd the H0 diagonal (GPU) and
    e2 = sigma . v_pt.  Returns (e2, TPSCIstate holding v_pt).  open_method
    selects open_matvec_thread's method ("sparse" for first-order spaces too
    large to list by target Fock block)."""
    from .state import project_out_state
    eng = engine or default_engine()
    norms = np.linalg.norm(ci_vector.coeffs, axis=0)
    sig = open_matvec_thread(ci_vector, cluster_ops, clustered_ham, nbody=nbody, thresh=thresh_foi,
                             prescreen=prescreen, engine=eng, method=open_method)
    sig = project_out_state(sig, ci_vector)
    h0 = eng.local_operator(clustered_ham, H0)
    if E0 is None:
        E0 = compute_expectation_value(ci_vector, cluster_ops, h0, engine=eng)
    E0 = np.asarray(E0, np.float64).reshape(-1)
    hd = compute_diagonal(sig, cluster_ops, h0, engine=eng) if len(sig) else np.zeros(0)
    sv = sig.coeffs
    v_pt = sv / (E0[None, :] / (norms[None, :] ** 2) - hd[:, None])
    e2 = np.einsum("ir,ir->r", sv, v_pt)
    sig.coeffs = v_pt
    if verbose:
        print(" PT1: %d configurations, e2 = %s" % (len(sig), np.array2string(e2, precision=8)))
    return e2, sig


def compute_qdpt_energy(ci_vector_in, cluster_ops, clustered_ham, nbody=4, H0="Hcmf", E0=None,
                        thresh_foi=1e-8, prescreen=True, threaded=True, verbose=1, engine=None,
                        open_method="dense"):
    """Quasi-degenerate second-order energies (compute_qdpt_energy,
    src/tpsci_pt2_energy.jl:311-361).

    As in the reference: the state is orthonormalised; E0 = <r|H|r> with the
    full Hamiltonian unless given; the first-order space X is the support of
    the screened open matvec at thresh_foi (open_matvec_thread) minus the
    state's own configurations (project_out!); with Hd the full-H diagonal
    over X (compute_diagonal) and H_ax the Hamiltonian between the model
    configurations a (those of the state) and X (build_full_H_parallel),
    the model-space matrix H_model + H_ax pinv(E0[1] I - diag(Hd)) H_ax^T is
    diagonalised, the denominator taking the first root's E0 for every
    configuration and the pseudo-inverse zeroing exact zeros.  Returns its
    eigenvalues, ascending (one per model configuration).

    On the GPU: the open matvec, Hd (fcg_space_diagonal, no plan), H_ax (a
    plan between X and the model space, applied to the identity) and
    H_model (build_full_H).  The n_model^2 reduction over X and the
    n_model-dimensional eigenproblem run on the host; the matrix is
    symmetrised before eigvalsh (the reference's product is symmetric up to
    rounding).  H0, prescreen and threaded are accepted for call
    compatibility (the reference prints H0 and uses it nowhere else).
    open_method selects open_matvec_thread's method."""
    from .state import project_out_state
    eng = engine or default_engine()
    ci = _copy_state(ci_vector_in)
    ci.set_vector(orthonormalize(ci.coeffs))
    if E0 is None:
        E0 = compute_expectation_value(ci, cluster_ops, clustered_ham, engine=eng)
    E0 = np.asarray(E0, np.float64).reshape(-1)
    sig = open_matvec_thread(ci, cluster_ops, clustered_ham, nbody=nbody, thresh=thresh_foi,
                             prescreen=prescreen, engine=eng, method=open_method)
    sig = project_out_state(sig, ci)
    n_model = len(ci)
    h_model = build_full_H(ci, cluster_ops, clustered_ham, engine=eng)
    if len(sig):
        hd = compute_diagonal(sig, cluster_ops, clustered_ham, engine=eng)
        pl = eng.pair_plan(sig, ci, cluster_ops, clustered_ham)
        h_xa = np.zeros((len(sig), n_model))
        for j0 in range(0, n_model, 16):
            j1 = min(n_model, j0 + 16)
            unit = np.zeros((n_model, j1 - j0))
            unit[np.arange(j0, j1), np.arange(j1 - j0)] = 1.0
            h_xa[:, j0:j1] = pl.sigma(unit)
        del pl
        den = E0[0] - hd
        dinv = np.zeros_like(den)
        nz = den != 0.0
        dinv[nz] = 1.0 / den[nz]
        h2 = (h_xa.T * dinv[None, :]) @ h_xa
    else:
        h2 = np.zeros((n_model, n_model))
    h_tot = h2 + h_model
    energies = np.linalg.eigvalsh(0.5 * (h_tot + h_tot.T))
    compute_qdpt_energy.last_stats = {"n_first_order": len(sig), "n_model": n_model}
    if verbose:
        print(" QDPT: %d first-order configurations, %d model configurations"
              % (len(sig), n_model))
        print(" %5s %12s %12s" % ("Root", "E(0)", "E(2)"))
        for r in range(min(len(E0), len(energies))):
            print(" %5d %12.8f %12.8f" % (r + 1, E0[r], energies[r]))
    return energies


def tpsci_ci(ci_vector, cluster_ops, clustered_ham, thresh_cipsi=1e-2, thresh_foi=1e-6,
             thresh_asci=None, thresh_var=None, thresh_spin=None, max_iter=10, conv_thresh=1e-4,
             nbody=4, incremental=True, ci_conv=1e-5, ci_max_iter=50, ci_max_ss_vecs=12,
             ci_lindep_thresh=1e-12, davidson=False, max_mem_ci=20.0, threaded=True, verbose=1,
             engine=None, open_method="dense"):
    """Selected CI in the TPS basis (tpsci_ci, src/tpsci.jl:39-363): each
    iteration adds the configurations of the clipped first-order
    wavefunction to the variational space, solves for R roots there
    (tps_ci_direct, or tps_ci_davidson when davidson=True or the dense H
    would exceed max_mem_ci GB), forms E0 = <H_cmf>, and builds the next
    first-order wavefunction from the variational vector (clipped at
    thresh_asci when given) -- incrementally as the reference does (the
    open matvec of the change of the vector since the last iteration, added
    to the last sigma after rotating it by the overlap) or, with
    incremental=False, by compute_pt1_wavefunction.  Stops when no root's
    energy changed by conv_thresh.  Every sigma, diagonal and solve runs on
    the GPU; the bookkeeping between them is keyed state algebra on the host.
    With thresh_spin, each iteration first adds the spin-completing
    configurations: the open matvec of the extract_S2 operator on the
    variational vector at threshold thresh_spin, its residual
    S^2|v> - |v><v|S^2|v> reported per root, and its configurations added
    with zero coefficients (src/tpsci.jl:126-161).  open_method selects
    open_matvec_thread's method ("sparse": generated per ket configuration
    and term, for first-order spaces too large to list by Fock block).
    Returns (e0, TPSCIstate of the variational vector); the last PT2 estimate
    is tpsci_ci.last_e2."""
    from .state import (add_states, clip_state, copy_state, overlap_states, project_out_state,
                        scale_state, zero_state, TPSCIstate)
    if len(ci_vector) == 0:
        raise ValueError("tpsci_ci: input vector has zero length")
    eng = engine or default_engine()
    R, N = ci_vector.R, ci_vector.N
    vec_var = copy_state(ci_vector)
    vec_pt = zero_state(ci_vector)
    e0 = np.zeros(R)
    e2 = np.zeros(R)
    e0_last = np.zeros(R)
    vec_asci_old = TPSCIstate(N, R)
    sig_old = TPSCIstate(N, R)
    h0 = eng.local_operator(clustered_ham, "Hcmf")
    callers_plans = set(eng._plans)
    for it in range(1, max_iter + 1):
        if it > 1:
            if thresh_var is not None:
                vec_var = clip_state(vec_var, thresh_var)
            vec_var = add_states(vec_var, zero_state(vec_pt))
        if thresh_spin is not None:
            # S^2|v> - |v><v|S^2|v>: its configurations spin-complete the space
            s2op = eng.s2_operator(cluster_ops)
            res = open_matvec_thread(vec_var, cluster_ops, s2op, nbody=nbody, thresh=thresh_spin,
                                     engine=eng, method=open_method)
            expval = overlap_states(vec_var, res)
            res = add_states(res, scale_state(vec_var, -expval))
            if verbose:
                for r in range(R):
                    print(" S^2 Residual %12.8f" % float(res.coeffs[:, r] @ res.coeffs[:, r]))
            n1 = len(vec_var)
            vec_var = add_states(vec_var, zero_state(res))
            if verbose:
                print(" Add spin completing states %8d -> %8d" % (n1, len(vec_var)))
        if 8.0 * len(vec_var) ** 2 * 1e-9 > max_mem_ci or davidson:
            vec_var.set_vector(orthonormalize(vec_var.coeffs))
            e0, vec_var = tps_ci_davidson(vec_var, cluster_ops, clustered_ham, conv_thresh=ci_conv,
                                          max_iter=ci_max_iter, max_ss_vecs=ci_max_ss_vecs,
                                          lindep_thresh=ci_lindep_thresh, engine=eng)
        else:
            e0, vec_var, _ = tps_ci_direct(vec_var, cluster_ops, clustered_ham, conv_thresh=ci_conv,
                                           max_ss_vecs=ci_max_ss_vecs, max_iter=ci_max_iter,
                                           lindep_thresh=ci_lindep_thresh, engine=eng)
        e0 = np.asarray(e0)
        efock = compute_expectation_value(vec_var, cluster_ops, h0, engine=eng)
        vec_asci = copy_state(vec_var)
        if thresh_asci is not None:
            vec_asci = clip_state(vec_asci, thresh_asci)
        if incremental:
            S = overlap_states(vec_asci_old, vec_asci)
            sig = scale_state(sig_old, S)
            del_v0 = add_states(vec_asci, scale_state(vec_asci_old, -S))
            vec_asci_old = copy_state(vec_asci)
            del_sig = open_matvec_thread(del_v0, cluster_ops, clustered_ham, nbody=nbody,
                                         thresh=thresh_foi, engine=eng, method=open_method)
            sig = add_states(sig, del_sig)
            sig_old = copy_state(sig)
            sig = project_out_state(sig, vec_asci)
            hd = compute_diagonal(sig, cluster_ops, h0, engine=eng) if len(sig) else np.zeros(0)
            norms = np.linalg.norm(vec_asci.coeffs, axis=0)
            sv = sig.coeffs
            v_pt = sv / (efock[None, :] / (norms[None, :] ** 2) - hd[:, None])
            e2 = np.einsum("ir,ir->r", sv, v_pt)
            vec_pt = copy_state(sig)
            vec_pt.coeffs = v_pt
        else:
            e2, vec_pt = compute_pt1_wavefunction(vec_asci, cluster_ops, clustered_ham, E0=efock,
                                                  thresh_foi=thresh_foi, nbody=nbody, verbose=0,
                                                  engine=eng, open_method=open_method)
        vec_pt = clip_state(vec_pt, thresh_cipsi)
        done = np.abs(e0_last - e0).max() < conv_thresh
        if verbose:
            print(" TPSCI iter %2d: dim %8d  E0 %s  E0+E2 %s%s"
                  % (it, len(vec_var), np.array2string(e0, precision=10),
                     np.array2string(e0 + e2, precision=10), "  *converged" if done else ""))
        eng.drop_plans(keep=callers_plans)  # this iteration's spaces are not used again
        if done:
            break
        e0_last = e0.copy()
    tpsci_ci.last_e2 = e2
    return e0, vec_var


def _copy_state(s):
    from .state import TPSCIstate
    o = TPSCIstate(s.N, s.R)
    o.fock = s.fock.copy()
    o.fock_start = s.fock_start.copy()
    o.cfg = s.cfg.copy()
    o.coeffs = s.coeffs.copy()
    return o

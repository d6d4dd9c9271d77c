"""Block-sparse Tucker outer loop (block_sparse_tucker, src/bst.jl:370-642)
and the Tucker algebra it runs on (src/hosvd.jl, src/type_BSTstate.jl).

The loop alternates, per iteration:
  compress the reference state (HOSVD of every block, singular values above
  thresh_var), solve in it, build the compressed first-order space
  (build_compressed_1st_order_state), its first-order wavefunction with the
  Hcmf zeroth-order Hamiltonian (compute_pt1_wavefunction), add that to the
  reference state, compress at thresh_pt, project the last solution in as
  the start vector, and solve again, until the energy moves by less than
  tol_tucker.

Every Hamiltonian application (sigma between two Tucker spaces, the
expectation values, the Davidson solves) runs in libfcg_hip.so through the
engine's BST pair plans and ci_solve.  The Tucker bookkeeping between them
(HOSVD truncation, non-orthogonal addition, projection between bases) is
host work on tensors of at most a few thousand elements, as in the
reference.

One difference of method: the reference forms the first-order space term by
term (form_sigma_block_expand, tucker_form_sigma_block_expand.jl), compresses
each term's block at thresh and adds them (nonorth_add, compress); here the
exact sigma of each target block (every term at once, on the GPU) is
compressed once at thresh, which keeps at least what the term-wise
truncation keeps.
"""
from collections import OrderedDict

import numpy as np

from .bst import BSTstate


# ---------------------------------------------------------------- Tucker algebra

def _unfold(A, i):
    return np.moveaxis(A, i, 0).reshape(A.shape[i], -1)


def _mode(A, M, i):
    """A with mode i contracted against the second index of M:
    out[.., J, ..] = sum_j M[J, j] A[.., j, ..]."""
    return np.moveaxis(np.tensordot(M, A, axes=([1], [i])), 0, i)


def transform(A, mats):
    """A x_1 mats[0] x_2 mats[1] ... (each mode contracted with the second
    index of its matrix); transform_basis(v, M', trans=false) of
    src/hosvd.jl:730-760 with M' = mats transposed."""
    for i, M in enumerate(mats):
        A = _mode(A, M, i)
    return A


def tucker_decompose(cores, thresh):
    """HOSVD of R core tensors sharing their factors (tucker_decompose,
    src/hosvd.jl:450-543): per mode, the eigenvectors of the summed Gram
    matrices of the unfoldings, ordered by |eigenvalue| descending, kept
    while sqrt|eigenvalue| exceeds thresh (none kept gives an empty mode).
    Returns (cores in the kept basis, factors)."""
    dims = cores[0].shape
    factors = []
    for i in range(len(dims)):
        G = np.zeros((dims[i], dims[i]))
        for A in cores:
            X = _unfold(A, i)
            G += X @ X.T
        w, U = np.linalg.eigh(G)
        w = np.abs(w)
        order = np.argsort(-w, kind="stable")
        s = np.sqrt(w[order])
        nkeep = int(np.count_nonzero(s > thresh))
        factors.append(U[:, order[:nkeep]])
    return [transform(A, [F.T for F in factors]) for A in cores], factors


class Tucker:
    """R core tensors (ranks) and per-cluster factors (range x rank)."""

    __slots__ = ("cores", "factors")

    def __init__(self, cores, factors):
        self.cores = cores
        self.factors = factors

    def ranks(self):
        return self.cores[0].shape

    def __len__(self):
        return int(np.prod(self.cores[0].shape))

    def norm(self):
        # norm(A::Tucker) = norm(A.core), the core tuple's 2-norm
        return float(np.sqrt(sum(np.sum(A * A) for A in self.cores)))

    def copy(self):
        return Tucker([A.copy() for A in self.cores], [F.copy() for F in self.factors])


def compress_tucker(t, thresh):
    """compress(::Tucker) (src/hosvd.jl:337-349): HOSVD of the core, the new
    factors being the old ones times the core's."""
    if len(t) == 0:
        return t
    cores, fs = tucker_decompose(t.cores, thresh)
    return Tucker(cores, [F @ G for F, G in zip(t.factors, fs)])


def nonorth_add_tuckers(tucks, thresh=1e-10):
    """nonorth_add(::Vector{Tucker}) (src/hosvd.jl:83-156): sorted by norm
    (largest first), per cluster an orthonormal basis of the concatenated
    factors (left singular vectors above thresh), and the cores transformed
    into it and summed."""
    order = sorted(range(len(tucks)), key=lambda k: -tucks[k].norm())
    tucks = [tucks[k] for k in order]
    if len(tucks) == 1:
        return tucks[0]
    N = len(tucks[0].factors)
    new_f = []
    for i in range(N):
        U, S, _ = np.linalg.svd(np.hstack([t.factors[i] for t in tucks]), full_matrices=False)
        new_f.append(U[:, :int(np.count_nonzero(S > thresh))])
    R = len(tucks[0].cores)
    cores = [np.zeros([F.shape[1] for F in new_f]) for _ in range(R)]
    for t in tucks:
        mats = [new_f[i].T @ t.factors[i] for i in range(N)]
        for r in range(R):
            cores[r] += transform(t.cores[r], mats)
    return Tucker(cores, new_f)


# ---------------------------------------------------------------- states

class TState:
    """OrderedDict fock -> OrderedDict ranges -> Tucker (the reference's
    BSTstate data, type_BSTstate.jl:13-18), ranges a tuple of 1-based
    inclusive (lo, hi) per cluster."""

    def __init__(self, N, R):
        self.N = N
        self.R = R
        self.data = OrderedDict()

    def blocks(self):
        for f, d in self.data.items():
            for rg, t in d.items():
                yield f, rg, t

    def set(self, f, rg, t):
        self.data.setdefault(f, OrderedDict())[rg] = t

    def get(self, f, rg):
        d = self.data.get(f)
        return None if d is None else d.get(rg)

    def __len__(self):
        return sum(len(t) for _, _, t in self.blocks())

    def copy(self):
        o = TState(self.N, self.R)
        for f, rg, t in self.blocks():
            o.set(f, rg, t.copy())
        return o

    def get_vector(self):
        cols = [np.concatenate([t.cores[r].ravel(order="F") for _, _, t in self.blocks()])
                if len(self) else np.zeros(0) for r in range(self.R)]
        return np.stack(cols, axis=1)

    def set_vector(self, v):
        v = np.asarray(v, np.float64)
        if v.ndim == 1:
            v = v[:, None]
        self.R = v.shape[1]
        at = 0
        for _, _, t in self.blocks():
            n = len(t)
            shp = t.ranks()
            t.cores = [v[at:at + n, r].reshape(shp, order="F").copy() for r in range(self.R)]
            at += n

    def to_bst(self):
        """The device's BSTstate (identity factors passed as such)."""
        st = BSTstate(self.N, self.R)
        for f, rg, t in self.blocks():
            facs = []
            for (lo, hi), F in zip(rg, t.factors):
                n = hi - lo + 1
                ident = F.shape == (n, n) and np.array_equal(F, np.eye(n))
                facs.append(None if ident else np.asfortranarray(F))
            st.add_block(f, list(rg), facs)
        st.coeffs = self.get_vector()
        return st

    @classmethod
    def from_bst(cls, st):
        o = cls(st.N, st.R)
        off = st.var_offsets()
        for b, (fi, rgs, facs) in enumerate(st.blocks):
            ranks = st.ranks(b)
            fs = [np.eye(hi - lo + 1) if F is None else np.asarray(F) for (lo, hi), F in
                  zip(rgs, facs)]
            cores = [st.coeffs[off[b]:off[b + 1], r].reshape(ranks, order="F").copy()
                     for r in range(st.R)]
            o.set(st.focks[fi], tuple(tuple(x) for x in rgs), Tucker(cores, fs))
        return o


def compress(v, thresh):
    """compress(::BSTstate) (type_BSTstate.jl:260-276): every block
    compressed, empty ones dropped."""
    o = TState(v.N, v.R)
    for f, rg, t in v.blocks():
        c = compress_tucker(t, thresh)
        if len(c) > 0:
            o.set(f, rg, c)
    return o


def compress_iteratively(v, thresh, maxiter=20):
    """compress_iteratively (type_BSTstate.jl:283-296)."""
    for _ in range(maxiter):
        n1 = len(v)
        v = compress(v, thresh)
        if not len(v) < n1:
            break
    return v


def nonorth_add(v1, v2):
    """nonorth_add! (type_BSTstate.jl:326-347): blocks of v2 added to v1's,
    a shared block through nonorth_add of the pair at 1e-7
    (src/hosvd.jl:297)."""
    o = v1.copy()
    for f, rg, t in v2.blocks():
        t1 = o.get(f, rg)
        o.set(f, rg, t.copy() if t1 is None else nonorth_add_tuckers([t1, t], thresh=1e-7))
    return o


def orthonormalize(v):
    """orthonormalize! (type_BSTstate.jl:476-491): Gram-Schmidt over roots."""
    x = v.get_vector()
    x[:, 0] /= np.linalg.norm(x[:, 0])
    for r in range(1, x.shape[1]):
        for r0 in range(r):
            x[:, r] -= x[:, r0] * (x[:, r0] @ x[:, r])
        x[:, r] /= np.linalg.norm(x[:, r])
    o = v.copy()
    o.set_vector(x)
    return o


def project_into_new_basis(v1, v2):
    """project_into_new_basis (type_BSTstate.jl:952-980): v1's cores carried
    into the factors of v2's blocks that v1 also holds; other blocks zero."""
    out = v2.copy()
    for f, rg, t in out.blocks():
        t1 = v1.get(f, rg)
        if t1 is None:
            t.cores = [np.zeros_like(A) for A in t.cores]
            continue
        S = [F1.T @ F2 for F1, F2 in zip(t1.factors, t.factors)]
        t.cores = [transform(A, [M.T for M in S]) for A in t1.cores]
    return out


def orth_dot(v1, v2):
    """orth_dot (type_BSTstate.jl:756-776): per root, the dot product of the
    cores (same blocks and factors)."""
    return np.einsum("ir,ir->r", v1.get_vector(), v2.get_vector())


def pq_spaces(bases, ref_fock):
    """The P space (the ground state of the reference sector) and its
    complement Q per cluster and Fock sector (BSTstate(clusters, fconfig,
    cluster_bases), type_BSTstate.jl:82-133; get_ortho_compliment,
    build_local_quantities.jl:12-33)."""
    p, q = [], []
    for c, b in enumerate(bases):
        pc, qc = {}, {}
        for sec in b.sectors():
            M = b.n_states(sec)
            if tuple(sec) == tuple(ref_fock[c]):
                pc[tuple(sec)] = (1, 1)
                if M > 1:
                    qc[tuple(sec)] = (2, M)
            elif M > 0:
                qc[tuple(sec)] = (1, M)
        p.append(pc)
        q.append(qc)
    return p, q


def reference_state(N, ref_fock, R=1):
    """The CMF product state as a one-block BST state (P on every cluster)."""
    v = TState(N, R)
    cores = [np.ones([1] * N) for _ in range(R)]
    v.set(tuple(tuple(x) for x in ref_fock), tuple((1, 1) for _ in range(N)),
          Tucker(cores, [np.ones((1, 1)) for _ in range(N)]))
    return v


# ---------------------------------------------------------------- GPU-backed steps

def _engine(engine):
    from .tpsci import default_engine
    return engine or default_engine()


def sigma_into(bra, ket, cluster_ops, clustered_ham, nbody=4, engine=None):
    """bra's blocks and factors holding <bra|H|ket> c_ket (build_sigma! into
    a zeroed bra, tucker_inner.jl:165-245), through a pair plan on the GPU."""
    eng = _engine(engine)
    out = bra.copy()
    if len(bra) == 0 or len(ket) == 0:
        out.set_vector(np.zeros((len(bra), ket.R)))
        return out
    s = eng.bst_pair_sigma(bra.to_bst(), ket.to_bst(), cluster_ops, clustered_ham, nbody)
    out.set_vector(s)
    return out


def expectation_value(v, cluster_ops, clustered_ham, nbody=4, engine=None):
    """<r|H|r> per root (compute_expectation_value, tucker_outer.jl)."""
    return orth_dot(v, sigma_into(v, v, cluster_ops, clustered_ham, nbody=nbody, engine=engine))


def ci_solve(v, cluster_ops, clustered_ham, conv_thresh=1e-5, max_iter=40, max_ss_vecs=12,
             lindep_thresh=1e-10, nbody=4, engine=None):
    """ci_solve (tucker_outer.jl:68-215) on the GPU: Davidson with the
    pseudo-canonical Hcmf preconditioner; the state comes back in the
    rotated factors, as the reference's does."""
    eng = _engine(engine)
    e, st = eng.bst_ci_solve(orthonormalize(v).to_bst(), cluster_ops, clustered_ham,
                             conv_thresh=conv_thresh, max_iter=max_iter,
                             max_ss_vecs=max_ss_vecs, lindep_thresh=lindep_thresh, nbody=nbody)
    return np.asarray(e), TState.from_bst(st)


def build_compressed_1st_order_state(psi, cluster_ops, clustered_ham, pq, nbody=4, thresh=1e-7,
                                     engine=None):
    """The compressed first-order space of psi (build_compressed_1st_order_state,
    tucker_outer.jl:717-811).  Targets as the reference enumerates them:
    for every Fock block of psi and TransferConfig, the bra Fock block
    F + TransferConfig (electron counts within the clusters, every sector in
    ClusterOps), and for every term and block of psi the TuckerConfigs that
    replace each active cluster's range by the P or the Q range of its bra
    sector (:830-870).  Each target block is formed exactly, every term at
    once, in the identity basis of its ranges (a pair plan on the GPU), then
    compressed at thresh; the whole state is compressed again until its
    dimension stops falling (compress_iteratively)."""
    p, q = pq
    N = psi.N
    targets = OrderedDict()
    for fock_psi, blocks in psi.data.items():
        for tc, terms in clustered_ham:
            fs = tuple((fock_psi[c][0] + tc[c][0], fock_psi[c][1] + tc[c][1]) for c in range(N))
            # electron counts within the clusters: a sector ClusterOps holds
            if any((fs[c], fs[c]) not in cluster_ops[c]["H"] for c in range(N)):
                continue
            for term in terms:
                if term.nbody > nbody:
                    continue
                avail = []
                for c in term.clusters:
                    opts = []
                    if fs[c] in p[c]:
                        opts.append(p[c][fs[c]])
                    if fs[c] in q[c]:
                        opts.append(q[c][fs[c]])
                    avail.append(opts)
                for rg_psi in blocks:
                    for prod in _product(avail):
                        rg = list(rg_psi)
                        for a, c in enumerate(term.clusters):
                            rg[c] = prod[a]
                        targets.setdefault(fs, OrderedDict())[tuple(rg)] = True
    bra = TState(N, psi.R)
    for fs, rgs in targets.items():
        for rg in rgs:
            dims = [hi - lo + 1 for lo, hi in rg]
            bra.set(fs, rg, Tucker([np.zeros(dims) for _ in range(psi.R)],
                                   [np.eye(d) for d in dims]))
    sig = sigma_into(bra, psi, cluster_ops, clustered_ham, nbody=nbody, engine=engine)
    sig = compress(sig, thresh)
    return compress_iteratively(sig, thresh)


def _product(lists):
    out = [()]
    for opts in lists:
        out = [o + (x,) for o in out for x in opts]
    return out


def compute_pt1_wavefunction(sig_in, psi0, cluster_ops, clustered_ham, H0="Hcmf", nbody=4,
                             engine=None, verbose=0):
    """First-order wavefunction in the space of sig_in
    (compute_pt1_wavefunction, tucker_pt.jl:73-160 and :173-254):
    sig_in's factors rotated to the pseudo-canonical H0 basis
    (form_1body_operator_diagonal!, pseudo_canon), then
    psi1 = (<X|H|0> - <X|F|0> - (E0 - F0) <X|0>) / (F0 - F_X + 1e-12)
    with E0 = <0|H|0>, F0 = <0|F|0>, F the clusters' H0, and
    E2 = E0 + <numerator|psi1>.  Returns (psi1, E2)."""
    from .bst import form_1body_operator_diagonal
    eng = _engine(engine)
    h0 = eng.local_operator(clustered_ham, H0)
    E0 = expectation_value(psi0, cluster_ops, clustered_ham, nbody=nbody, engine=eng)
    F0 = expectation_value(psi0, cluster_ops, h0, nbody=nbody, engine=eng)
    z = sig_in.copy()
    z.set_vector(np.zeros((len(z), psi0.R)))
    rot, fdiag = form_1body_operator_diagonal(z.to_bst(), cluster_ops, H0=H0, pseudo_canon=True)
    basis = TState.from_bst(rot)
    Sx = project_into_new_basis(psi0, basis)
    num = sigma_into(basis, psi0, cluster_ops, clustered_ham, nbody=nbody, engine=eng)
    xf0 = sigma_into(basis, psi0, cluster_ops, h0, nbody=nbody, engine=eng)
    x = num.get_vector() - xf0.get_vector() - Sx.get_vector() * (E0 - F0)[None, :]
    psi1v = x / (F0[None, :] - fdiag[:, None] + 1e-12)
    num.set_vector(x)
    psi1 = basis.copy()
    psi1.set_vector(psi1v)
    ecorr = np.einsum("ir,ir->r", x, psi1v)
    E2 = E0 + ecorr
    if verbose:
        for r in range(len(E2)):
            print(" State %3d: E(0) %14.8f  E(PT2) corr %14.8f  E(2) %14.8f"
                  % (r + 1, E0[r], ecorr[r], E2[r]))
    return psi1, E2


def block_sparse_tucker(input_vec, cluster_ops, clustered_ham, pq, max_iter=20, nbody=4,
                        H0="Hcmf", thresh_var=1e-4, thresh_foi=1e-6, thresh_pt=1e-5,
                        thresh_spin=None, ci_conv=1e-5, ci_max_iter=50, ci_max_ss_vecs=12,
                        ci_lindep_thresh=1e-10, resolve_ss=False, do_pt=True, tol_tucker=1e-6,
                        solver="davidson", verbose=1, engine=None, foi_builder=None):
    """The BST outer loop (block_sparse_tucker, src/bst.jl:370-642), same
    arguments and order of steps; pq = pq_spaces(bases, reference Fock
    configuration) are the P / Q spaces of the reference's BSTstate.
    Returns (e_var, state); the per-iteration energies and dimensions are
    kept as block_sparse_tucker.history.  foi_builder replaces
    build_compressed_1st_order_state (same arguments), e.g. with a
    restatement that builds the space term by term."""
    if thresh_spin is not None:
        raise NotImplementedError("block_sparse_tucker: the S^2 spin extension (thresh_spin) "
                                  "is not built")
    eng = _engine(engine)
    ref_vec = input_vec.copy()
    var_vec_old = input_vec.copy()
    e_last = None
    e_var = None
    hist = []
    converged = False
    for it in range(max_iter):
        ref_vec = orthonormalize(compress(ref_vec, thresh_var))
        if resolve_ss:
            e0, ref_vec = ci_solve(ref_vec, cluster_ops, clustered_ham, conv_thresh=ci_conv,
                                   max_iter=ci_max_iter, max_ss_vecs=ci_max_ss_vecs,
                                   lindep_thresh=ci_lindep_thresh, nbody=nbody, engine=eng)
        else:
            e0 = expectation_value(ref_vec, cluster_ops, clustered_ham, nbody=nbody, engine=eng)
        if it == 0:
            e_last = e0
        pt1 = (foi_builder or build_compressed_1st_order_state)(
            ref_vec, cluster_ops, clustered_ham, pq, nbody=nbody, thresh=thresh_foi, engine=eng)
        e_pt2 = None
        if do_pt:
            pt1, e_pt2 = compute_pt1_wavefunction(pt1, ref_vec, cluster_ops, clustered_ham,
                                                  H0=H0, nbody=nbody, engine=eng)
        if len(pt1) == 0:
            raise RuntimeError("block_sparse_tucker: empty first-order space")
        dim_foi = len(pt1)
        pt1 = compress(pt1, thresh_foi)
        var_vec = orthonormalize(compress(nonorth_add(ref_vec, pt1), thresh_pt))
        var_vec = orthonormalize(project_into_new_basis(var_vec_old, var_vec))
        e_var, var_vec = ci_solve(var_vec, cluster_ops, clustered_ham, conv_thresh=ci_conv,
                                  max_iter=ci_max_iter, max_ss_vecs=ci_max_ss_vecs,
                                  lindep_thresh=ci_lindep_thresh, nbody=nbody, engine=eng)
        var_vec_old = var_vec.copy()
        hist.append({"iter": it + 1, "e_ref": np.asarray(e0).copy(), "dim_ref": len(ref_vec),
                     "dim_foi": dim_foi, "dim_foi_compressed": len(pt1),
                     "e_pt2": None if e_pt2 is None else np.asarray(e_pt2).copy(),
                     "e_var": np.asarray(e_var).copy(), "dim_var": len(var_vec)})
        if verbose:
            h = hist[-1]
            print(" BST iter %3d  E(ref) %s  dim %d  FOIS %d -> %d  E(PT2) %s  E(BST) %s  dim %d"
                  % (h["iter"], np.array2string(h["e_ref"], precision=8), h["dim_ref"],
                     h["dim_foi"], h["dim_foi_compressed"],
                     "-" if h["e_pt2"] is None else np.array2string(h["e_pt2"], precision=8),
                     np.array2string(h["e_var"], precision=8), h["dim_var"]), flush=True)
        ref_vec = var_vec
        if np.max(np.abs(np.asarray(e_last) - np.asarray(e_var))) < tol_tucker:
            converged = True
            break
        e_last = e_var
    block_sparse_tucker.history = hist
    block_sparse_tucker.converged = converged
    return np.asarray(e_var), ref_vec

"""Term-wise restatement of the BST first-order space (test infrastructure
only; numpy on the host).

The reference builds the compressed first-order space of a BST state term
by term (build_compressed_1st_order_state and
_build_compressed_1st_order_state_job, src/tucker_outer.jl:717-900;
form_sigma_block_expand, src/tucker_form_sigma_block_expand.jl):

  for every bra Fock block F + TransferConfig, every term of that
  TransferConfig and every ket Tucker block, the bra TuckerConfigs that put
  the P or the Q range of the bra sector on each active cluster (Julia's
  Iterators.product, first active cluster fastest); for each, the term's
  block expanded on its active clusters:
    g_a[p, I, l] = sum_J Γ_a[p, I, J] U_a[J, l]   (ket factor U_a)
    1-body: op[I, l] = H[I, J] U[J, l];
    3- and 4-body terms first cut each g_a to the left singular vectors of
      its (I, p l) matricisation with s * max(1, max_r |c_r|) * |V| > thresh
      (precontract), the kept vectors becoming that cluster's factor;
    core[I_A, l_rest] = sign * sum_{p, l_A} V[p] prod_a g_a[p_a, I_a, l_a]
                        * c[l_A, l_rest],
  the other clusters keeping the ket's factors; each such block compressed
  at thresh, empty ones dropped; per bra TuckerConfig the blocks added
  (nonorth_add, 1e-10) and compressed again; the state compressed until
  its dimension stops falling.

It is the reference's construction, against which the GPU build's exact
construction (every term summed, compressed once) is compared: run through
the same outer loop it reproduces the reference's block_sparse_tucker
energy on h12 (test/test_bst.jl:87).
"""
from collections import OrderedDict

import numpy as np

from fermicg_amd.bst_loop import (TState, Tucker, compress_iteratively, compress_tucker,
                                  nonorth_add_tuckers)

_LET = "abcdefghijklmnopqrstuvwxyz"


def _sign(term, fock_ket):
    """compute_terms_state_sign (src/type_ClusteredTerm.jl:181-197)."""
    s = 0
    for c, par in zip(term.clusters, term.parity):
        if par % 2 == 1:
            s ^= sum(fock_ket[x][0] + fock_ket[x][1] for x in range(c)) & 1
    return -1.0 if s else 1.0


def expand_term_block(term, cluster_ops, fock_b, rg_b, fock_k, rg_k, tk, thresh):
    """form_sigma_block_expand for one term, bra TuckerConfig and ket block."""
    A = list(term.clusters)
    sign = _sign(term, fock_k)
    factors = list(tk.factors)
    if term.nbody == 1:
        c = A[0]
        (lb, hb), (lk, hk) = rg_b[c], rg_k[c]
        H = np.asarray(cluster_ops[c][term.ops[0]][(fock_b[c], fock_k[c])])
        op = H[lb - 1:hb, lk - 1:hk] @ tk.factors[c]
        cores = [sign * np.moveaxis(np.tensordot(op, C, axes=([1], [c])), 0, c) for C in tk.cores]
        factors[c] = np.eye(hb - lb + 1)
        return Tucker(cores, factors)
    gs = []
    for a, c in enumerate(A):
        (lb, hb), (lk, hk) = rg_b[c], rg_k[c]
        G = np.asarray(cluster_ops[c][term.ops[a]][(fock_b[c], fock_k[c])])[:, lb - 1:hb, lk - 1:hk]
        gs.append(np.einsum("pIJ,Jl->pIl", G, tk.factors[c]))
    V = np.asarray(term.ints)
    newf = [np.eye(g.shape[1]) for g in gs]
    if term.nbody >= 3:
        scale0 = max(1.0, max(float(np.linalg.norm(C)) for C in tk.cores))
        scale1 = scale0 * float(np.linalg.norm(V))
        for a, g in enumerate(gs):
            npp, nI, nl = g.shape
            D = g.transpose(1, 0, 2).reshape(nI, npp * nl, order="F")
            U, S, Vt = np.linalg.svd(D, full_matrices=False)
            nkeep = int(np.count_nonzero(S * scale1 > thresh))
            newf[a] = U[:, :nkeep]
            gk = (S[:nkeep, None] * Vt[:nkeep, :]).reshape(nkeep, npp, nl, order="F")
            gs[a] = gk.transpose(1, 0, 2)
    k = term.nbody
    p_ix, I_ix, l_ix = _LET[:k], _LET[k:2 * k], _LET[2 * k:3 * k]
    spec = p_ix + "," + ",".join(p_ix[a] + I_ix[a] + l_ix[a] for a in range(k)) + "->" + I_ix + l_ix
    op = np.einsum(spec, V, *gs, optimize=True)          # [I_A..., l_A...]
    cores = []
    for C in tk.cores:
        Cm = np.moveaxis(C, A, list(range(k)))            # [l_A..., rest...]
        out = np.tensordot(op, Cm, axes=(list(range(k, 2 * k)), list(range(k))))
        cores.append(sign * np.moveaxis(out, list(range(k)), A))
    for a, c in enumerate(A):
        factors[c] = newf[a]
    return Tucker(cores, factors)


def build_compressed_1st_order_state_termwise(psi, cluster_ops, clustered_ham, pq, nbody=4,
                                              thresh=1e-7, engine=None):
    """The reference's build_compressed_1st_order_state, term by term (see
    the module docstring).  `engine` is accepted for the outer loop's call
    and not used."""
    p, q = pq
    N = psi.N
    jobs = OrderedDict()
    for fock_psi, blocks in psi.data.items():
        for tc, terms in clustered_ham:
            fs = tuple((fock_psi[c][0] + tc[c][0], fock_psi[c][1] + tc[c][1]) for c in range(N))
            if any((fs[c], fs[c]) not in cluster_ops[c]["H"] for c in range(N)):
                continue
            jobs.setdefault(fs, []).append((terms, fock_psi, blocks))
    out = TState(N, psi.R)
    for fs, joblist in jobs.items():
        data = OrderedDict()
        for terms, fock_psi, blocks in joblist:
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
                for rg_psi, tk in blocks.items():
                    for prod in _product_first_fastest(avail):
                        rg = list(rg_psi)
                        for a, c in enumerate(term.clusters):
                            rg[c] = prod[a]
                        t = expand_term_block(term, cluster_ops, fs, rg, fock_psi, rg_psi, tk,
                                              thresh)
                        if len(t) == 0:
                            continue
                        t = compress_tucker(t, thresh)
                        if len(t) == 0:
                            continue
                        data.setdefault(tuple(rg), []).append(t)
        for rg, tucks in data.items():
            out.set(fs, rg, compress_tucker(nonorth_add_tuckers(tucks, thresh=1e-10), thresh))
    return compress_iteratively(out, thresh)


def _product_first_fastest(lists):
    """Iterators.product order: the first list varies fastest."""
    out = [()]
    for opts in lists:
        out = [o + (x,) for x in opts for o in out]
    return out

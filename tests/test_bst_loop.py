"""The BST outer loop's host algebra and the loop itself on the CPU oracle
engine (block_sparse_tucker, src/bst.jl:370-642; Tucker algebra of
src/hosvd.jl and src/type_BSTstate.jl).

Pins by another route than the code under test:
* tucker_decompose at thresh 0 reconstructs the tensor; truncating one mode
  loses exactly the squares of the dropped singular values of that
  unfolding (numpy SVD), and keeps those above thresh;
* nonorth_add of Tuckers with different factors reconstructs to the sum of
  their dense tensors; project_into_new_basis equals the dense projection
  onto the new factors;
* from the CMF state of the reference's h12 case, with nothing truncated,
  the BST first-order energy equals E_CMF plus the TPSCI PT2 of the single
  CMF configuration (1e-10 Eh);
* the reference's own case (test/test_bst.jl:87, 1e-6): with the
  first-order space built term by term as the reference builds it
  (oracle/bst_termwise.py), the loop reaches its energy;
* at the reference's thresholds the BST energy lies between FCI (dense
  eigenvalue of the oracle's build_full_H over every Fock block of a
  complete-basis 4-cluster system) and CMF.
"""
import numpy as np
import pytest

from fermicg_amd.bst_loop import (Tucker, block_sparse_tucker, compress_tucker, nonorth_add_tuckers,
                                  pq_spaces, project_into_new_basis, reference_state, transform,
                                  tucker_decompose, TState)
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from oracle.pyoracle import Oracle
from tests.oracle_engine import OracleEngine
from tests.pt2_cases import four_clusters


def _dense(t, r=0):
    return transform(t.cores[r], t.factors)


def test_tucker_decompose_exact_and_truncated():
    rng = np.random.default_rng(1)
    A = [rng.normal(size=(4, 5, 3)), rng.normal(size=(4, 5, 3))]
    cores, fs = tucker_decompose(A, 0.0)
    for r in range(2):
        assert np.allclose(transform(cores[r], fs), A[r], atol=1e-12)
    # a tensor of multilinear rank (2, 5, 3): mode 0 truncates at thresh
    B = [np.einsum("ia,ajk->ijk", rng.normal(size=(4, 2)), rng.normal(size=(2, 5, 3)))]
    B[0] += 1e-4 * rng.normal(size=B[0].shape)
    s = np.linalg.svd(B[0].reshape(4, -1), compute_uv=False)
    t = compress_tucker(Tucker(B, [np.eye(4), np.eye(5), np.eye(3)]), 1e-2)
    assert t.ranks()[0] == int(np.sum(s > 1e-2)) == 2
    err = np.linalg.norm(_dense(t) - B[0]) ** 2
    assert abs(err - np.sum(s[2:] ** 2)) < 1e-10 + 1e-6 * err


def test_nonorth_add_and_projection():
    rng = np.random.default_rng(2)
    dims = (5, 4, 3)

    def rand_tucker(ranks):
        fs = [np.linalg.qr(rng.normal(size=(d, k)))[0] for d, k in zip(dims, ranks)]
        return Tucker([rng.normal(size=ranks)], fs)

    ts = [rand_tucker((2, 2, 1)), rand_tucker((3, 1, 2)), rand_tucker((1, 2, 2))]
    s = nonorth_add_tuckers(ts, thresh=1e-10)
    assert np.allclose(_dense(s), sum(_dense(t) for t in ts), atol=1e-12)
    v1, v2 = TState(3, 1), TState(3, 1)
    rg = ((1, 5), (1, 4), (1, 3))
    v1.set(((1, 0),) * 3, rg, ts[0])
    v2.set(((1, 0),) * 3, rg, ts[1])
    p = project_into_new_basis(v1, v2).get(((1, 0),) * 3, rg)
    F = ts[1].factors
    ref = transform(_dense(ts[0]), [f.T for f in F])
    assert np.allclose(p.cores[0], ref, atol=1e-12)


def _fci(case):
    full = TPSCIstate.full_blocks(case.bases, case.all_focks, R=1)
    H = Oracle(flatten_ops(case.ops, [len(c) for c in case.clusters]), flatten_ham(case.H),
               full).full_H()
    return np.linalg.eigvalsh(H)[0]


def test_bst_pt1_equals_tpsci_pt2_from_cmf(oracle_lib):
    """h12's cluster bases are eigenstates of the embedded (Hcmf) cluster
    Hamiltonians, so from the CMF product state the BST first-order energy
    with nothing truncated (compute_pt1_wavefunction over
    build_compressed_1st_order_state at thresh -1, which keeps every
    direction of each target block, so the pseudo-canonical H0 is exact
    there) is E_CMF plus the
    Epstein-Nesbet sum over every other configuration with Hcmf
    denominators: the TPSCI PT2 of the single CMF configuration (the
    oracle's orc_tpsci_pt2, pinned to test/test_tpsci.jl's h12 values)."""
    from fermicg_amd.bst_loop import build_compressed_1st_order_state, compute_pt1_wavefunction
    from fermicg_amd.cluster_ops import OP_NAMES
    from fermicg_amd.operator import extract_1body_operator
    from tests.h12_cases import h12_problem
    p, _ = h12_problem()
    eng = OracleEngine(nthreads=8)
    pq = pq_spaces(p.bases, p.ref)
    v = reference_state(5, p.ref)
    foi = build_compressed_1st_order_state(v, p.ops, p.H, pq, thresh=-1.0, engine=eng)
    psi1, e2_bst = compute_pt1_wavefunction(foi, v, p.ops, p.H, engine=eng)
    cmf = TPSCIstate.from_blocks(5, [(p.ref, [(1,) * 5])], R=1)
    cmf.set_vector(np.ones((1, 1)))
    orc = Oracle(p.fops, p.fham, cmf)
    e_cmf = orc.full_H()[0, 0]
    h0 = Oracle(p.fops, flatten_ham(extract_1body_operator(p.H, "Hcmf")), cmf).full_H()[0, 0]
    e2, _ = orc.pt2(cmf.coeffs, np.array([h0]), h0_op=OP_NAMES.index("Hcmf"), thresh=0.0,
                    nthreads=8)
    print("FOIS", len(foi), "BST E(PT2)", e2_bst[0], "CMF + TPSCI e2", e_cmf + e2[0], flush=True)
    assert abs(e2_bst[0] - (e_cmf + e2[0])) < 1e-10


def test_bst_loop_variational_bounds(oracle_lib):
    case = four_clusters(seed=3, R=1)
    pq = pq_spaces(case.bases, case.ref)
    v = reference_state(4, case.ref)
    e, st = block_sparse_tucker(v, case.ops, case.H, pq, max_iter=10, thresh_var=1e-2,
                                thresh_foi=1e-3, thresh_pt=np.sqrt(1e-5), ci_conv=1e-8,
                                resolve_ss=True, tol_tucker=1e-6, verbose=0,
                                engine=OracleEngine(nthreads=8))
    hist = block_sparse_tucker.history
    e_cmf = hist[0]["e_ref"][0]
    efci = _fci(case)
    print("e_cmf", e_cmf, "e_bst", e[0], "fci", efci, [h["dim_var"] for h in hist], flush=True)
    assert efci - 1e-10 <= e[0] <= e_cmf + 1e-10
    assert hist[-1]["dim_var"] < 784


def test_split_block_covers_variables_in_order():
    """Blocks above 64 variables are cut into boxes whose embeddings,
    concatenated, are the block's own (same columns, same order)."""
    from fermicg_amd.bst import BOX_MAX, _split_block
    rng = np.random.default_rng(5)

    def emb(rg, fa, N):
        K = None
        for c in range(N):
            lo, hi = rg[c]
            f = fa[c]
            E = np.zeros((12, f.shape[1] if f is not None else hi - lo + 1))
            E[lo - 1:hi, :] = f if f is not None else np.eye(hi - lo + 1)
            K = E if K is None else np.kron(E, K)
        return K

    for _ in range(100):
        N = int(rng.integers(1, 5))
        rgs, facs = [], []
        for c in range(N):
            lo, n = int(rng.integers(1, 3)), int(rng.integers(1, 10))
            rgs.append((lo, lo + n - 1))
            facs.append(None if rng.random() < 0.5 else rng.normal(size=(n, int(rng.integers(1, n + 1)))))
        boxes = _split_block(rgs, facs)
        for r, f in boxes:
            assert np.prod([(x.shape[1] if x is not None else h - l + 1)
                            for (l, h), x in zip(r, f)]) <= BOX_MAX
        assert np.array_equal(np.hstack([emb(r, f, N) for r, f in boxes]), emb(rgs, facs, N))


@pytest.mark.slow
def test_termwise_loop_reproduces_reference_h12(oracle_lib):
    """The reference's BST test (test/test_bst.jl:72-87): block_sparse_tucker
    on h12 with its arguments, the first-order space built term by term as
    the reference builds it (oracle/bst_termwise.py), every other step the
    loop of fermicg_amd.bst_loop on the oracle engine: the energy within the
    test's 1e-6 (it lands 3e-8 from -18.329455496968325)."""
    from oracle.bst_termwise import build_compressed_1st_order_state_termwise
    from tests.bst_loop_cases import E_VAR_REF, H12_ARGS
    from tests.h12_cases import h12_problem
    p, _ = h12_problem()
    e, _ = block_sparse_tucker(reference_state(5, p.ref), p.ops, p.H, pq_spaces(p.bases, p.ref),
                               engine=OracleEngine(nthreads=8), verbose=1,
                               foi_builder=build_compressed_1st_order_state_termwise, **H12_ARGS)
    print("e_var", e[0], "reference", E_VAR_REF, "diff", e[0] - E_VAR_REF, flush=True)
    assert block_sparse_tucker.converged
    assert abs(e[0] - E_VAR_REF) < 1e-6

"""GPU parity of the TPSCI sigma build through the C ABI.

Oracle: the C restatement of tps_ci_matvec / build_full_H / compute_diagonal
(oracle/fcg_oracle.c, pinned by tests/test_producer_pins.py).  Bars: FP64
sigma within 1e-12 relative (max-norm), diagonal exact to 1e-12, FP32 within
1e-5 and split-BF16 within 1e-4 relative, Davidson energies within 1e-10 Eh
of dense diagonalisation of the oracle's H.
"""
import numpy as np
import pytest

from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import Engine, cluster_sizes_of, tps_ci_davidson, tps_ci_matvec, compute_diagonal
from fermicg_amd.workloads import tpsci_small
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


@pytest.fixture(scope="module")
def small():
    p, st = tpsci_small(seed=0, R=3)
    orc = Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                 flatten_ham(p.clustered_ham), st)
    return p, st, orc


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_sigma_fp64_full_blocks(eng, small):
    p, st, orc = small
    s_gpu = tps_ci_matvec(st, p.cluster_ops, p.clustered_ham, engine=eng)
    s_ref = orc.sigma(st.coeffs)
    print("n_cfg", len(st), "rel", rel(s_gpu, s_ref), flush=True)
    assert s_gpu.shape == s_ref.shape
    assert rel(s_gpu, s_ref) < 1e-12


def test_diagonal(eng, small):
    p, st, orc = small
    hd = compute_diagonal(st, p.cluster_ops, p.clustered_ham, engine=eng)
    assert np.abs(hd - orc.diagonal()).max() < 1e-12 * np.abs(hd).max()
    # the plan's stored diagonal (fcg_diagonal, used by Davidson) agrees
    hp = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham).diagonal()
    assert np.abs(hp - hd).max() < 1e-12 * np.abs(hd).max()


# BF16 keeps 8 mantissa bits (unit roundoff 2^-9 = 2e-3 per element): 1e-2,
# also for fp32_bf16c (c rounded to bf16 in the one-row apply)
@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16x2", 1e-4), ("bf16", 1e-2),
                                      ("fp32_bf16c", 1e-2)])
def test_sigma_reduced_precision(eng, small, prec, tol):
    p, st, orc = small
    s_gpu = tps_ci_matvec(st, p.cluster_ops, p.clustered_ham, precision=prec, engine=eng)
    assert rel(s_gpu, orc.sigma(st.coeffs)) < tol


def test_sigma_sparse_blocks(eng, small):
    # drop a seeded third of the configurations: Fock blocks are no longer
    # full products, so the library falls back to single-configuration boxes
    p, st, orc = small
    rng = np.random.Generator(np.random.PCG64(7))
    keep = rng.random(len(st)) > 0.33
    blocks = []
    for fi, a, b in st.blocks():
        confs = [tuple(int(x) for x in st.cfg[i]) for i in range(a, b) if keep[i]]
        if confs:
            blocks.append((st.focks()[fi], confs))
    sp = TPSCIstate.from_blocks(st.N, blocks, R=2)
    sp.set_vector(rng.normal(size=(len(sp), 2)))
    s_gpu = tps_ci_matvec(sp, p.cluster_ops, p.clustered_ham, engine=eng)
    orc2 = Oracle(orc_fops(p), flatten_ham(p.clustered_ham), sp)
    assert rel(s_gpu, orc2.sigma(sp.coeffs)) < 1e-12


def orc_fops(p):
    return flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops))


def test_davidson_matches_dense(eng, small):
    p, st, orc = small
    H = orc.full_H()
    w = np.linalg.eigvalsh(H)[:3]
    e, v = tps_ci_davidson(st, p.cluster_ops, p.clustered_ham, conv_thresh=1e-9, max_iter=300,
                           engine=eng)
    print("davidson", e, w, v.davidson_stats, flush=True)
    assert np.abs(e - w).max() < 1e-10
    r = H @ v.coeffs - v.coeffs * e
    assert np.abs(r).max() < 1e-8


def test_davidson_mixed_precision(eng, small):
    p, st, orc = small
    w = np.linalg.eigvalsh(orc.full_H())[:3]
    for prec in ("fp32", "bf16x2", "bf16", "fp32_bf16c"):
        e, v = tps_ci_davidson(st, p.cluster_ops, p.clustered_ham, conv_thresh=1e-8,
                               max_iter=400, precision=prec, refresh_k=4, engine=eng)
        print(prec, e - w, v.davidson_stats, flush=True)
        assert np.abs(e - w).max() < 1e-6  # |dE| <= 1 uEh against FP64


def test_davidson_large_subspace_host_eigensolver(eng, small):
    """A subspace of up to 40 x 3 = 120 vectors: past 96 the projected
    eigenproblem leaves the GPU eigensolver for the host one; the energies
    still match dense diagonalisation, without a preconditioner so that the
    subspace does grow past 96 before convergence."""
    p, st, orc = small
    H = orc.full_H()
    w = np.linalg.eigvalsh(H)[:3]
    e, v = tps_ci_davidson(st, p.cluster_ops, p.clustered_ham, conv_thresh=1e-9, max_iter=400,
                           max_ss_vecs=40, precond=False, engine=eng)
    print("davidson", e, w, v.davidson_stats, flush=True)
    assert v.davidson_stats["iterations"] > 32   # m = 3 + 3 (it - 1) passed 96
    assert np.abs(e - w).max() < 1e-10

"""GPU parity of the device-joined sigma for sparse TPSCI spaces (every Fock
block holds a subset of its configurations, so every row is one
configuration): the rank-1 configuration sets of the large-cluster workloads
and a thinned full space.

Oracle: the C restatement of tps_ci_matvec / compute_diagonal / build_full_H
(oracle/fcg_oracle.c).  Bars: FP64 sigma 1e-12 relative (max-norm), diagonal
1e-12 relative, FP32 1e-5, BF16X2 1e-4 and BF16 1e-2, Davidson energies
1e-10 Eh of dense diagonalisation of the oracle's H (FP64 and mixed, whose
convergence is decided on FP64 residuals).
"""
import numpy as np
import pytest

from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import Engine, cluster_sizes_of, compute_diagonal, tps_ci_davidson, tps_ci_matvec
from fermicg_amd.workloads import build_problem, rank1_space, tpsci_small
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


def _oracle(p, st, nbody=4):
    return Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                  flatten_ham(p.clustered_ham, nbody), st)


def _thinned_small():
    p, st = tpsci_small(seed=0, R=2)
    rng = np.random.Generator(np.random.PCG64(11))
    blocks = []
    for fi, a, b in st.blocks():
        idx = [i for i in range(a, b) if rng.random() > 0.3]
        if b - a > 1 and len(idx) == b - a:
            idx = idx[1:]            # no Fock block stays whole: one-configuration rows only
        if not idx:
            idx = [a]
        blocks.append((st.focks()[fi], [tuple(int(x) for x in st.cfg[i]) for i in idx]))
    sp = TPSCIstate.from_blocks(3, blocks, R=2)
    sp.set_vector(rng.normal(size=(len(sp), 2)))
    return p, sp


def _rank1(n_clusters, M, transfer, frac, seed, R):
    p = build_problem(4 * n_clusters, n_clusters, M, 8 * n_clusters, seed, max_transfer=transfer)
    st = rank1_space(p, R=R, frac=frac, seed=seed)
    st.set_vector(np.random.Generator(np.random.PCG64(seed)).normal(size=(len(st), R)))
    return p, st


def _thinned_large_blocks():
    """3 clusters, M=8: Fock blocks of up to 512 configurations, thinned so
    no block stays whole but several keep more than 256 -- beyond the device
    join's limit, so the plan takes the host join with one-configuration
    boxes."""
    p = build_problem(12, 3, 8, 24, seed=0, max_transfer=1)
    st = TPSCIstate.full_blocks(p.bases, p.focks, R=2)
    rng = np.random.Generator(np.random.PCG64(5))
    blocks = []
    for fi, a, b in st.blocks():
        idx = [i for i in range(a, b) if rng.random() > 0.25]
        if b - a > 1 and len(idx) == b - a:
            idx = idx[1:]
        if not idx:
            idx = [a]
        blocks.append((st.focks()[fi], [tuple(int(x) for x in st.cfg[i]) for i in idx]))
    sp = TPSCIstate.from_blocks(3, blocks, R=2)
    sp.set_vector(rng.normal(size=(len(sp), 2)))
    assert np.diff(sp.fock_start).max() > 256
    return p, sp


CASES = {
    "thinned_3c_M8_over256": _thinned_large_blocks,
    "thinned_3c": _thinned_small,
    "rank1_4c_M8": lambda: _rank1(4, 8, 2, 0.5, 5, 3),
    "rank1_5c_M16_full": lambda: _rank1(5, 16, 1, 1.0, 9, 2),   # blocks of up to 76 rows
    "rank1_6c_M16": lambda: _rank1(6, 16, 2, 0.159, 7, 1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_sparse_sigma_fp64(eng, name):
    p, st = CASES[name]()
    print("case", name, "rows", len(st), flush=True)
    s = tps_ci_matvec(st, p.cluster_ops, p.clustered_ham, engine=eng)
    orc = _oracle(p, st)
    ref = orc.sigma(st.coeffs)
    rel = np.abs(s - ref).max() / np.abs(ref).max()
    print("rel", rel, flush=True)
    assert rel < 1e-12
    hd = compute_diagonal(st, p.cluster_ops, p.clustered_ham, engine=eng)
    assert np.abs(hd - orc.diagonal()).max() < 1e-12 * np.abs(hd).max()
    hp = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham).diagonal()
    assert np.abs(hp - hd).max() < 1e-12 * np.abs(hd).max()


# BF16 keeps 8 mantissa bits (unit roundoff 2^-9 = 2e-3 per element): 1e-2,
# also for fp32_bf16c, whose one-row apply gathers c rounded to bf16
@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16x2", 1e-4), ("bf16", 1e-2),
                                      ("fp32_bf16c", 1e-2)])
def test_sparse_sigma_reduced(eng, prec, tol):
    p, st = CASES["rank1_4c_M8"]()
    s = tps_ci_matvec(st, p.cluster_ops, p.clustered_ham, precision=prec, engine=eng)
    ref = _oracle(p, st).sigma(st.coeffs)
    assert np.abs(s - ref).max() / np.abs(ref).max() < tol


def test_sparse_nbody2(eng):
    p, st = CASES["thinned_3c"]()
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham, nbody=2)
    s = pl.sigma(st.coeffs)
    ref = _oracle(p, st, nbody=2).sigma(st.coeffs)
    assert np.abs(s - ref).max() / np.abs(ref).max() < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16", "fp32_bf16c"])
def test_sparse_davidson(eng, prec):
    p, st = CASES["rank1_5c_M16_full"]()
    H = _oracle(p, st).full_H()
    ref = np.linalg.eigvalsh(H)[:2]
    e, v = tps_ci_davidson(st, p.cluster_ops, p.clustered_ham, conv_thresh=1e-8, max_iter=300,
                           precision=prec, engine=eng)
    print(prec, e, ref, v.davidson_stats, flush=True)
    assert np.abs(e - ref).max() < 1e-10

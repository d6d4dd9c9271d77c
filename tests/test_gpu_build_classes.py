"""The block build's LDS classes build the same blocks.

Each (bra box, ket box) pair is built by the smallest LDS class that holds it
(tiny, small, middle, full).  fcg_set_build_class_cap forces every pair of a
plan into a class at most as small as the cap, so the same pairs run
through each class's kernel.  Cases:
  * a TPSCI space of whole Fock blocks (identity factors: the separable
    Kronecker form) and a thinned one;
  * BST spaces (Tucker factors: factor compression and inactive-cluster
    overlaps);
  * PT2 (fused sigma; the tiny class is never used there).
Bars: FP64 sigma bit for bit across caps and within 1e-12 relative of the
oracle (oracle/fcg_oracle.c via oracle/pyoracle.py); PT2 e2 within 1e-12
relative across caps (its global atomics order the last bits) and 1e-11 of
the oracle's orc_tpsci_pt2.
"""
import numpy as np
import pytest

from fermicg_amd.bst import bst_pq_state, build_sigma, embedding
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import Engine, cluster_sizes_of, compute_pt2_energy
from fermicg_amd.workloads import build_problem, tpsci_small
from oracle.pyoracle import Oracle
from tests.pt2_cases import three_clusters_m4
from tests.test_gpu_pt2 import _oracle_pt2

pytestmark = pytest.mark.gpu

CAPS = (3, 2, 1, 0)  # tiny (default), small, middle, full


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.ctx.set_build_class_cap(3)
    e.clear()


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _with_cap(eng, cap):
    eng.ctx.set_build_class_cap(cap)
    eng.clear()  # plans are cached per space: build them again


@pytest.mark.parametrize("thin", [False, True])
def test_tpsci_classes(eng, thin):
    p, st = tpsci_small(seed=0, R=2)
    if thin:  # every third configuration dropped: boxes of ragged sizes
        blocks = []
        focks = st.focks()
        for fi, a, b in st.blocks():
            confs = [tuple(int(x) for x in st.cfg[i]) for i in range(a, b) if i % 3 != 1]
            if len(confs) > 1:
                blocks.append((focks[fi], confs))
        st = TPSCIstate.from_blocks(st.N, blocks, R=2)
    st.set_vector(np.random.Generator(np.random.PCG64(5)).normal(size=(len(st), 2)))
    ref = Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                 flatten_ham(p.clustered_ham), st).sigma(st.coeffs)
    out = {}
    for cap in CAPS:
        _with_cap(eng, cap)
        out[cap] = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham).sigma(st.coeffs)
        print("thin", thin, "cap", cap, "rel", _rel(out[cap], ref), flush=True)
        assert _rel(out[cap], ref) < 1e-12
    for cap in CAPS[1:]:
        assert np.array_equal(out[cap], out[CAPS[0]])


@pytest.mark.parametrize("case", [(12, 3, 6, 2, 3), (8, 4, 4, 2, 2)])
def test_bst_classes(eng, case):
    no, nc, M, mq, mr = case
    p = build_problem(no, nc, M, 2 * no, seed=3, max_transfer=2)
    bst = bst_pq_state(p.bases, p.focks, p.ref_fock, max_q=mq, max_rank=mr, R=2, seed=4)
    tps = TPSCIstate.full_blocks(p.bases, p.focks)
    U = embedding(bst, p.bases, tps)
    Hb = U.T @ Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                      flatten_ham(p.clustered_ham), tps).full_H() @ U
    ref = Hb @ bst.coeffs
    out = {}
    for cap in CAPS:
        _with_cap(eng, cap)
        sig = bst.copy_structure()
        build_sigma(sig, bst, p.cluster_ops, p.clustered_ham, engine=eng)
        out[cap] = sig.coeffs.copy()
        print(case, "cap", cap, "rel", _rel(out[cap], ref), flush=True)
        assert _rel(out[cap], ref) < 1e-12
    for cap in CAPS[1:]:
        assert np.array_equal(out[cap], out[CAPS[0]])


def test_pt2_classes(eng):
    case = three_clusters_m4()
    _, e2_ref = _oracle_pt2(case)
    out = {}
    for cap in CAPS:
        _with_cap(eng, cap)
        out[cap] = compute_pt2_energy(case.state, case.ops, case.H, thresh_foi=0.0, verbose=0,
                                      engine=eng)
        print("pt2 cap", cap, out[cap], "oracle", e2_ref, flush=True)
        assert np.abs(out[cap] - e2_ref).max() < 1e-11 * np.abs(e2_ref).max()
    for cap in CAPS[1:]:
        assert np.abs(out[cap] - out[CAPS[0]]).max() < 1e-12 * np.abs(e2_ref).max()

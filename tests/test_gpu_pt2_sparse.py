"""GPU parity of the sparse first-order-space PT2 (fcg_tpsci_pt2_sparse).

Oracle: orc_tpsci_pt2, the C restatement of compute_pt2_energy with the
_collect_significant_thread! element test (pinned by the dense closed form
and, through tpsci_ci, by the reference's own h12 values), on a 6-cluster
case at several thresholds and on the 3- and 4-cluster cases; the dense GPU
PT2 on the same inputs; the reference's h12 e0 + e2 (test/test_tpsci.jl:
123-131) through compute_pt2_energy with the sparse method.  Bars: e2 within
1e-11 relative of the oracle; repeated calls bit-identical.
"""
import numpy as np
import pytest

from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.tpsci import Engine, compute_pt2_energy
from oracle.pyoracle import Oracle
from tests.pt2_cases import four_clusters, six_clusters_rank1, three_clusters_m4
from tests.test_pt2_oracle import _e0

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


@pytest.fixture(scope="module")
def six():
    return six_clusters_rank1()


def _oracle(case, e0, thresh, nbody=4):
    st = case.state
    fo = flatten_ops(case.ops, [len(c) for c in case.clusters])
    e2, _ = Oracle(fo, flatten_ham(case.H, nbody), st).pt2(st.coeffs, e0, nbody=nbody,
                                                           thresh=thresh)
    return e2


@pytest.mark.parametrize("thresh", [0.0, 1e-6, 1e-4, 1e-3])
def test_sparse_pt2_six_clusters(eng, six, thresh):
    e0 = _e0(six, six.state)
    ref = _oracle(six, e0, thresh)
    e2 = compute_pt2_energy(six.state, six.ops, six.H, E0=e0, thresh_foi=thresh, verbose=0,
                            method="sparse", engine=eng)
    st = compute_pt2_energy.last_stats
    print(thresh, e2, ref, st, flush=True)
    assert np.abs(e2 - ref).max() < 1e-11 * np.abs(ref).max()
    again = compute_pt2_energy(six.state, six.ops, six.H, E0=e0, thresh_foi=thresh, verbose=0,
                               method="sparse", engine=eng)
    assert np.array_equal(again, e2)   # fixed-order accumulation


def test_sparse_equals_dense_six_clusters(eng, six):
    e0 = _e0(six, six.state)
    kw = dict(E0=e0, thresh_foi=1e-5, verbose=0, engine=eng)
    d = compute_pt2_energy(six.state, six.ops, six.H, method="dense", **kw)
    s = compute_pt2_energy(six.state, six.ops, six.H, method="sparse", **kw)
    assert np.abs(s - d).max() < 1e-11 * np.abs(d).max()


@pytest.mark.parametrize("name,nbody,thresh", [("four", 4, 0.0), ("four", 2, 0.0),
                                               ("three", 4, 0.0), ("three", 4, 1e-4)])
def test_sparse_pt2_small_cases(eng, name, nbody, thresh):
    case = {"four": four_clusters, "three": three_clusters_m4}[name]()
    e0 = _e0(case, case.state)
    ref = _oracle(case, e0, thresh, nbody)
    e2 = compute_pt2_energy(case.state, case.ops, case.H, nbody=nbody, E0=e0, thresh_foi=thresh,
                            verbose=0, method="sparse", engine=eng)
    print(name, nbody, thresh, e2, ref, flush=True)
    assert np.abs(e2 - ref).max() < 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("chunk", [1, 997])
def test_sparse_pt2_chunked_accumulation(eng, six, chunk, monkeypatch):
    """Accumulation cut into many chunks of whole target Fock blocks (chunk=1
    puts every target block in a chunk of its own) agrees with the oracle and
    keeps the contribution and configuration counts of the unchunked call."""
    e0 = _e0(six, six.state)
    thresh = 1e-6
    ref = _oracle(six, e0, thresh)
    kw = dict(E0=e0, thresh_foi=thresh, verbose=0, method="sparse", engine=eng)
    whole = compute_pt2_energy(six.state, six.ops, six.H, **kw)
    st0 = dict(compute_pt2_energy.last_stats)
    monkeypatch.setenv("FCG_PT2_CHUNK", str(chunk))
    e2 = compute_pt2_energy(six.state, six.ops, six.H, **kw)
    st = compute_pt2_energy.last_stats
    print(chunk, e2, whole, ref, st["n_chunks"], st0["n_chunks"], flush=True)
    assert st0["n_chunks"] == 1 and st["n_chunks"] > 1
    for k in ("n_contributions", "n_configs", "n_first_order", "n_items_kept", "n_kets_kept"):
        assert st[k] == st0[k], k
    assert np.abs(e2 - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(e2 - whole).max() < 1e-13 * np.abs(whole).max()


@pytest.mark.parametrize("M,nnz,thresh", [(4, 6, 1e-3), (4, 6, 1e-5), (8, 10, 1e-5)])
def test_sparse_pt2_twelve_clusters(eng, M, nnz, thresh):
    """The 12-cluster, 48-orbital shape against the oracle's restatement over
    the configurations reached (orc_tpsci_pt2_sparse, pinned to the dense
    target-block restatement on the small cases): e2 to 1e-11 relative and
    the number of first-order configurations exactly (1.2e8 at M=8)."""
    from tests.pt2_cases import twelve_clusters_rank1
    case = twelve_clusters_rank1(M=M, nnz=nnz)
    st = case.state
    e0 = np.array([-30000.0])   # below every H0 diagonal: no small denominators
    fo = flatten_ops(case.ops, [len(c) for c in case.clusters])
    ref, _, nx = Oracle(fo, flatten_ham(case.H, 4), st).pt2_sparse(st.coeffs, e0, thresh=thresh)
    e2 = compute_pt2_energy(st, case.ops, case.H, E0=e0, thresh_foi=thresh, verbose=0,
                            method="sparse", engine=eng)
    stats = compute_pt2_energy.last_stats
    print(M, nnz, thresh, e2, ref, nx, stats, flush=True)
    assert stats["n_first_order"] == nx
    assert np.abs(e2 - ref).max() < 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("case_name,thresh", [("six", 1e-6), ("twelve", 1e-5)])
def test_sparse_pt2_fp32_contraction(eng, six, case_name, thresh):
    """The mixed form (FP32 chain and elements, FP64 tests, values and sums)
    against the FP64 engine and the oracle: |e2(FP32) - e2(FP64)| within
    1e-5 of |e2| (FP32 rounding of the contractions; elements within that
    rounding of the threshold may flip), the kept-contribution count within
    1e-4 of FP64's, and repeated calls bit-identical."""
    from tests.pt2_cases import twelve_clusters_rank1
    case = six if case_name == "six" else twelve_clusters_rank1(M=8, nnz=10)
    st = case.state
    e0 = _e0(case, st) if case_name == "six" else np.array([-30000.0])
    kw = dict(E0=e0, thresh_foi=thresh, verbose=0, method="sparse", engine=eng)
    e64 = compute_pt2_energy(st, case.ops, case.H, **kw)
    s64 = dict(compute_pt2_energy.last_stats)
    e32 = compute_pt2_energy(st, case.ops, case.H, precision="fp32", **kw)
    s32 = dict(compute_pt2_energy.last_stats)
    again = compute_pt2_energy(st, case.ops, case.H, precision="fp32", **kw)
    print(case_name, e64, e32, np.abs(e32 - e64) / np.abs(e64), s64["n_contributions"],
          s32["n_contributions"], flush=True)
    assert np.array_equal(again, e32)
    assert np.abs(e32 - e64).max() < 1e-5 * np.abs(e64).max()
    assert abs(s32["n_contributions"] - s64["n_contributions"]) <= 1e-4 * s64["n_contributions"]


def test_sparse_pt2_boxed_and_unboxed_space(eng, six):
    """A space already cut into boxes by a plan and a freshly uploaded one
    (read directly) give the same e2 and statistics bit for bit."""
    from fermicg_amd import _lib
    from fermicg_amd.cluster_ops import OP_NAMES
    st = six.state
    e0 = _e0(six, st)
    ham, ops = eng.ham(six.H, 4), eng.ops(six.ops)
    fresh = _lib.DeviceSpace.tpsci(eng.ctx, st)
    boxed = _lib.DeviceSpace.tpsci(eng.ctx, st)
    pl = _lib.Plan.create(eng.ctx, ham, ops, boxed)   # forms the boxes
    del pl
    kw = dict(nbody=4, h0_op=OP_NAMES.index("Hcmf"), thresh=1e-6)
    e_fresh, s_fresh = _lib.tpsci_pt2_sparse(eng.ctx, ham, ops, fresh, st.coeffs, e0, **kw)
    e_boxed, s_boxed = _lib.tpsci_pt2_sparse(eng.ctx, ham, ops, boxed, st.coeffs, e0, **kw)
    print(e_fresh, e_boxed, flush=True)
    assert np.array_equal(e_fresh, e_boxed)
    for k in ("n_items", "n_contributions", "n_configs", "n_first_order"):
        assert s_fresh[k] == s_boxed[k], k

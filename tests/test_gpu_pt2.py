"""GPU parity of the open (first-order space) sigma and of PT2 through the C ABI.

Oracle: orc_tpsci_pt2 (the restatement of compute_pt2_energy at
thresh_foi = 0, pinned to the dense formula in test_pt2_oracle.py) and the
oracle's build_full_H over every Fock block.  Bars: e2 and E0 within 1e-11
relative; the open sigma within 1e-12 relative (max-norm).
"""
import numpy as np
import pytest

from fermicg_amd import _lib
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import Engine, compute_pt2_energy
from oracle.pyoracle import Oracle
from tests.pt2_cases import four_clusters, three_clusters_m4
from tests.test_pt2_oracle import _e0

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


def _oracle_pt2(case, nbody=4, thresh=0.0):
    st = case.state
    fo = flatten_ops(case.ops, [len(c) for c in case.clusters])
    e0 = _e0(case, st)
    e2, _ = Oracle(fo, flatten_ham(case.H, nbody), st).pt2(st.coeffs, e0, nbody=nbody,
                                                          thresh=thresh)
    return e0, e2


@pytest.mark.parametrize("name,nbody", [("four", 4), ("four", 2), ("three", 4), ("three_full", 4)])
def test_pt2_parity(eng, name, nbody):
    case = {"four": lambda: four_clusters(),
            "three": lambda: three_clusters_m4(),
            "three_full": lambda: three_clusters_m4(R=1, full_blocks=True)}[name]()
    print("case", name, "nbody", nbody, "n_var", len(case.state), flush=True)
    e0_ref, e2_ref = _oracle_pt2(case, nbody)
    e2 = compute_pt2_energy(case.state, case.ops, case.H, nbody=nbody, thresh_foi=0.0, verbose=0,
                            engine=eng)
    st = compute_pt2_energy.last_stats
    print("e2", e2, "oracle", e2_ref, "stats", st, flush=True)
    assert np.abs(e2 - e2_ref).max() < 1e-11 * np.abs(e2_ref).max()
    # E0 computed on the GPU (Hcmf-only operator) against the oracle's
    e2b = compute_pt2_energy(case.state, case.ops, case.H, nbody=nbody, E0=e0_ref, thresh_foi=0.0,
                             verbose=0, engine=eng)
    assert np.abs(e2b - e2).max() < 1e-11 * np.abs(e2_ref).max()


@pytest.mark.parametrize("thresh", [1e-9, 1e-6, 1e-4, 1e-2])
def test_pt2_screened(eng, thresh):
    # the reference's thresh_foi screening: GPU against the oracle's
    # restatement of _collect_significant_thread! at the same threshold
    case = three_clusters_m4()
    e0_ref, e2_ref = _oracle_pt2(case, thresh=thresh)
    _, e2_exact = _oracle_pt2(case)
    e2 = compute_pt2_energy(case.state, case.ops, case.H, thresh_foi=thresh, verbose=0,
                            engine=eng)
    print("thresh", thresh, "gpu", e2, "oracle", e2_ref, "exact", e2_exact, flush=True)
    assert np.abs(e2 - e2_ref).max() < 1e-11 * np.abs(e2_ref).max()
    if thresh >= 1e-4:   # the screening changes the answer: the test sees it
        assert np.abs(e2_ref - e2_exact).max() > 1e-9 * np.abs(e2_exact).max()


def test_open_sigma_pair_plan(eng):
    # sigma over every configuration of every Fock block from a ket space:
    # the pair plan against the dense H of the whole space
    case = four_clusters(seed=4, R=2)
    st = case.state
    full = TPSCIstate.full_blocks(case.bases, case.all_focks, R=1)
    fo = flatten_ops(case.ops, [len(c) for c in case.clusters])
    H = Oracle(fo, flatten_ham(case.H), full).full_H()
    focks = full.focks()
    index = {(focks[fi], tuple(int(x) for x in full.cfg[i])): i
             for fi, a, b in full.blocks() for i in range(a, b)}
    vf = st.focks()
    var_idx = np.array([index[(vf[fi], tuple(int(x) for x in st.cfg[i]))]
                        for fi, a, b in st.blocks() for i in range(a, b)])
    ref = H[:, var_idx] @ st.coeffs
    bra = _lib.DeviceSpace.tpsci(eng.ctx, full)
    ket = _lib.DeviceSpace.tpsci(eng.ctx, st)
    pl = _lib.Plan.create_pair(eng.ctx, eng.ham(case.H), eng.ops(case.ops), bra, ket)
    s = pl.sigma(st.coeffs)
    print("pair plan", pl.stats(), flush=True)
    assert s.shape == ref.shape
    assert np.abs(s - ref).max() < 1e-12 * np.abs(ref).max()


def test_fused_sigma_fixed_order(eng):
    # the fused sigma (PT2's first-order sigma and the open sigma) sums each
    # bra row over its block pairs in pair order, with no float atomics: PT2
    # energies and the open sigma are bit-identical from run to run, and the
    # open sigma agrees with the stored plan's sigma
    from fermicg_amd.workloads import pt2_medium
    p, st = pt2_medium(M=7)
    print("pt2_medium M=7 n_var", len(st), flush=True)
    e2s = [compute_pt2_energy(st, p.cluster_ops, p.clustered_ham, thresh_foi=1e-9, verbose=0,
                              engine=eng) for _ in range(3)]
    print("e2", e2s[0], flush=True)
    for e in e2s[1:]:
        assert np.array_equal(e, e2s[0])
    sp = _lib.DeviceSpace.tpsci(eng.ctx, st)
    ham, ops = eng.ham(p.clustered_ham), eng.ops(p.cluster_ops)
    sig = [_lib.open_sigma(eng.ctx, ham, ops, sp, sp, st.coeffs)[0] for _ in range(3)]
    for s in sig[1:]:
        assert np.array_equal(s, sig[0])
    ref = _lib.Plan.create(eng.ctx, ham, ops, sp).sigma(st.coeffs)
    assert np.abs(sig[0] - ref).max() < 1e-12 * np.abs(ref).max()

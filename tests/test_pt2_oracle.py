"""Pins of the oracle's restatement of compute_pt2_energy (CPU only).

The restatement (oracle/fcg_oracle.c orc_tpsci_pt2: jobs per target Fock
block, per-term open matvec, project_out!, H0 diagonal, _sum_pt2) is checked
against what the mathematics fixes, computed by a different route:

  e2[r] = sum_{x not in V} ((H c)[x, r])^2 / (E0[r] - H0_xx)

with H the dense Hamiltonian over every Fock block of the problem (the
oracle's build_full_H, itself pinned to determinant FCI in
test_producer_pins.py) and H0_xx read straight from the "Hcmf" blocks.
A dropped term, a wrong sign, a wrong target-block set or a missing
projection all change this number.  Bar: 1e-11 relative.
"""
import numpy as np
import pytest

from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.operator import extract_1body_operator
from fermicg_amd.state import TPSCIstate, copy_state
from oracle.pyoracle import Oracle
from tests.pt2_cases import four_clusters, three_clusters_m4


def _orc(case, state, nbody=4, ham=None):
    fops = flatten_ops(case.ops, [len(c) for c in case.clusters])
    return Oracle(fops, flatten_ham(ham if ham is not None else case.H, nbody), state)


def _e0(case, st):
    h0 = extract_1body_operator(case.H, op_string="Hcmf")
    H0 = _orc(case, st, ham=h0).full_H()
    return np.einsum("ir,ij,jr->r", st.coeffs, H0, st.coeffs)


def _dense_pt2(case, st, e0, nbody=4):
    full = TPSCIstate.full_blocks(case.bases, case.all_focks, R=1)
    H = _orc(case, full, nbody=nbody).full_H()
    focks = full.focks()
    index = {}
    hd = np.zeros(len(full))
    for fi, a, b in full.blocks():
        f = focks[fi]
        for i in range(a, b):
            cfg = tuple(int(x) for x in full.cfg[i])
            index[(f, cfg)] = i
            hd[i] = sum(case.ops[ci]["Hcmf"][(f[ci], f[ci])][cfg[ci] - 1, cfg[ci] - 1]
                        for ci in range(len(f)))
    vf = st.focks()
    var_idx = np.array([index[(vf[fi], tuple(int(x) for x in st.cfg[i]))]
                        for fi, a, b in st.blocks() for i in range(a, b)])
    sig = H[:, var_idx] @ st.coeffs
    keep = np.ones(len(full), bool)
    keep[var_idx] = False
    return np.array([np.sum(sig[keep, r] ** 2 / (e0[r] - hd[keep])) for r in range(st.R)])


@pytest.mark.parametrize("nbody", [4, 2])
def test_pt2_matches_dense_four_clusters(oracle_lib, nbody):
    case = four_clusters()
    st = case.state
    e0 = _e0(case, st)
    e2, nj = _orc(case, st, nbody=nbody).pt2(st.coeffs, e0, nbody=nbody)
    ref = _dense_pt2(case, st, e0, nbody=nbody)
    print("jobs", nj, "e2", e2, "dense", ref)
    assert nj > 0
    assert np.abs(e2 - ref).max() < 1e-11 * np.abs(ref).max()


def test_pt2_matches_dense_three_clusters(oracle_lib):
    case = three_clusters_m4(R=1)
    st = case.state
    e0 = _e0(case, st)
    e2, _ = _orc(case, st).pt2(st.coeffs, e0)
    ref = _dense_pt2(case, st, e0)
    assert np.abs(e2 - ref).max() < 1e-11 * np.abs(ref).max()


def test_pt2_invariants(oracle_lib):
    case = four_clusters(seed=5, R=1)
    st = case.state
    orc = _orc(case, st)
    e0 = _e0(case, st)
    e2, _ = orc.pt2(st.coeffs, e0)
    # quadratic in the coefficients at zero threshold
    e2s, _ = orc.pt2(3.0 * st.coeffs, e0)
    assert abs(e2s[0] - 9.0 * e2[0]) < 1e-12 * abs(e2s[0])
    # a threshold above every element screens everything out
    assert orc.pt2(st.coeffs, e0, thresh=1e6)[0][0] == 0.0
    # the element threshold converges to the exact value
    d = [abs(orc.pt2(st.coeffs, e0, thresh=t)[0][0] - e2[0]) for t in (1e-3, 1e-6, 1e-14)]
    print("threshold errors", d)
    assert d[2] < 1e-12 * abs(e2[0]) and d[1] <= d[0]
    # the variational space equal to the whole space leaves nothing outside
    full = TPSCIstate.full_blocks(case.bases, case.all_focks, R=1)
    v = np.random.Generator(np.random.PCG64(1)).normal(size=(len(full), 1))
    v /= np.linalg.norm(v)
    assert _orc(case, full).pt2(v, [-1.0])[0][0] == 0.0


def _with_zeros(st, frac, seed):
    """The state with a seeded fraction of its configurations zeroed (the
    reference configuration of each root kept)."""
    rng = np.random.default_rng(seed)
    c = st.coeffs.copy()
    zero = rng.random(len(c)) < frac
    zero[np.abs(c).argmax(axis=0)] = False
    c[zero] = 0.0
    out = copy_state(st)
    out.coeffs = c
    return out


@pytest.mark.parametrize("name,nbody,thresh,zeros", [
    ("four", 4, 0.0, 0.0), ("four", 2, 1e-4, 0.5), ("three", 4, 0.0, 0.7),
    ("six", 4, 1e-5, 0.0), ("six", 4, 1e-3, 0.8)])
def test_pt2_sparse_restatement_matches_dense_targets(oracle_lib, name, nbody, thresh, zeros):
    """orc_tpsci_pt2_sparse (a map over the configurations reached, zero
    kets skipped) against orc_tpsci_pt2 (dense target blocks): the same
    jobs, terms and element test, so equal to rounding."""
    from tests.pt2_cases import six_clusters_rank1
    case = {"four": four_clusters, "three": three_clusters_m4, "six": six_clusters_rank1}[name]()
    st = case.state
    if zeros:
        st = _with_zeros(st, zeros, 3)
    e0 = _e0(case, st) if name != "six" else np.full(st.R, -30.0)
    orc = _orc(case, st, nbody=nbody)
    ref, nj = orc.pt2(st.coeffs, e0, nbody=nbody, thresh=thresh)
    e2, nj2, nx = orc.pt2_sparse(st.coeffs, e0, nbody=nbody, thresh=thresh)
    print(name, thresh, zeros, e2, ref, nj, nj2, nx)
    assert nx > 0 and nj2 <= nj
    assert np.abs(e2 - ref).max() < 1e-12 * np.abs(ref).max()

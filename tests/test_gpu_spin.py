"""Spin on the GPU: the S^2 operator (extract_S2 terms over the S+, S-, Sz,
S2 Γ) applied by libfcg_hip.so equals the oracle's, and the reference's h9
open-shell TPSCI values (test/test_openshell.jl:9-61: spin-adapted bases,
add_spin_focksectors, tpsci_ci with and without thresh_spin,
compute_pt2_energy) are reached with every sigma, open matvec, diagonal and
solve on the GPU, at the reference's 1e-8."""
import numpy as np
import pytest

from fermicg_amd.cluster_ops import compute_cluster_eigenbasis, compute_cluster_ops
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.operator import extract_ClusteredTerms, extract_S2
from fermicg_amd.state import TPSCIstate, fock_configs_within
from fermicg_amd.tpsci import Engine, compute_pt2_energy, compute_s2, tps_ci_matvec, tpsci_ci
from oracle.pyoracle import Oracle
from tests.h12_cases import h12_problem
from tests.spin_cases import E0_REF, E02_REF, E0_SPIN_REF, H9, PT2_THRESH, TPSCI_ARGS

pytestmark = pytest.mark.gpu


def test_gpu_s2_sigma_h12_1a2b(oracle_lib):
    _, f = h12_problem()
    ints, clusters = f["ints"], f["clusters"]
    bases = compute_cluster_eigenbasis(ints, clusters, max_roots=1000)
    ops = compute_cluster_ops(bases, ints)
    ref = ((1, 1), (0, 1), (0, 0), (0, 0), (0, 0))
    st = TPSCIstate.full_blocks(bases, fock_configs_within(bases, ref, 100), R=3)
    st.set_vector(np.random.default_rng(0).normal(size=(len(st), 3)))
    sizes = [len(c) for c in clusters]
    S2op = extract_S2(len(clusters), sizes)
    eng = Engine(0)
    s = tps_ci_matvec(st, ops, S2op, engine=eng)
    ref_s = Oracle(flatten_ops(ops, sizes), flatten_ham(S2op), st).sigma(st.coeffs)
    assert np.abs(s - ref_s).max() / np.abs(ref_s).max() < 1e-12
    # eigenvectors of H carry S(S+1) (1 alpha, 2 beta: doublets and quartets)
    Hm = Oracle(flatten_ops(ops, sizes), flatten_ham(extract_ClusteredTerms(ints, clusters)),
                st).full_H()
    S2m = Oracle(flatten_ops(ops, sizes), flatten_ham(S2op), st).full_H()
    w, v = np.linalg.eigh(Hm + 0.01 * S2m)
    st.set_vector(v[:, :6])
    s2 = compute_s2(st, ops, engine=eng)
    assert np.abs(s2 - np.array([0.75, 0.75, 3.75, 0.75, 0.75, 3.75])).max() < 1e-10


def test_gpu_h9_openshell_tpsci_pins():
    h = H9()
    eng = Engine(0)
    e0, v0 = tpsci_ci(h.start(), h.ops, h.H, engine=eng, verbose=0, **TPSCI_ARGS)
    print("e0", e0 - E0_REF, flush=True)
    assert np.abs(e0 - E0_REF).max() < 1e-8
    e2 = compute_pt2_energy(v0, h.ops, h.H, thresh_foi=PT2_THRESH, engine=eng, verbose=0)
    assert np.abs(e0 + e2 - E02_REF).max() < 1e-8
    e0s, v0s = tpsci_ci(h.start(), h.ops, h.H, engine=eng, verbose=0, thresh_spin=1e-5,
                        **TPSCI_ARGS)
    print("e0 spin", e0s - E0_SPIN_REF, flush=True)
    assert np.abs(e0s - E0_SPIN_REF).max() < 1e-8
    s2 = compute_s2(v0s, h.ops, engine=eng)
    print("s2", s2, flush=True)
    assert np.abs(s2 - np.array([0.75, 0.75, 3.75])).max() < 1e-6   # spin eigenstates

"""Spin operators, spin-adapted cluster bases and the S^2 extension of
tpsci_ci, pinned on the CPU by values the reference's tests hold.

* test/test_s2.jl:9-49: in 12 orbitals with 1 alpha and 2 beta electrons
  [H, S^2] = 0 and the lowest ten eigenstates of H + 0.01 S^2 have <S^2>
  0.75, 0.75, 3.75, 0.75, 0.75, 3.75, 0.75, 0.75, 0.75, 0.75 -- here in
  the tensor-product basis (complete cluster bases over every Fock
  configuration), with the S^2 of extract_S2 and the Γ of S+, S-, Sz, S2.
* test/test_openshell.jl:9-61: h9 with spin-adapted bases
  (compute_cluster_eigenbasis_spin), add_spin_focksectors, tpsci_ci and
  compute_pt2_energy: e0 and e0 + e2 of the reference at 1e-8, and e0 with
  thresh_spin = 1e-5.
* possible_spin_focksectors against brute-force enumeration.
"""
import itertools

import numpy as np
import pytest

from fermicg_amd.cluster_ops import compute_cluster_eigenbasis, compute_cluster_ops
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.operator import extract_ClusteredTerms, extract_S2
from fermicg_amd.state import TPSCIstate, fock_configs_within, possible_spin_focksectors
from fermicg_amd.tpsci import compute_pt2_energy, compute_s2, tpsci_ci
from oracle.pyoracle import Oracle
from tests.h12_cases import h12_problem
from tests.oracle_engine import OracleEngine
from tests.spin_cases import E0_REF, E02_REF, E0_SPIN_REF, H9, PT2_THRESH, TPSCI_ARGS

S2_REF = np.array([0.75, 0.75, 3.75, 0.75, 0.75, 3.75, 0.75, 0.75, 0.75, 0.75])


def test_s2_commutes_and_multiplets(oracle_lib):
    _, f = h12_problem()
    ints, clusters = f["ints"], f["clusters"]
    bases = compute_cluster_eigenbasis(ints, clusters, max_roots=1000)
    ops = compute_cluster_ops(bases, ints)
    ref = ((1, 1), (0, 1), (0, 0), (0, 0), (0, 0))   # 1 alpha, 2 beta electrons
    focks = fock_configs_within(bases, ref, 100)
    st = TPSCIstate.full_blocks(bases, focks)
    assert len(st) == 12 * 66                        # the whole determinant space
    sizes = [len(c) for c in clusters]
    fo = flatten_ops(ops, sizes)
    Hm = Oracle(fo, flatten_ham(extract_ClusteredTerms(ints, clusters)), st).full_H()
    S2 = Oracle(fo, flatten_ham(extract_S2(len(clusters), sizes)), st).full_H()
    assert np.abs(S2 - S2.T).max() < 1e-13
    assert np.linalg.norm(S2 @ Hm - Hm @ S2) < 1e-12
    w, v = np.linalg.eigh(Hm + 0.01 * S2)
    ss = np.einsum("ik,ij,jk->k", v[:, :10], S2, v[:, :10])
    assert np.abs(ss - S2_REF).max() < 1e-12


def test_spin_ladder_bases_are_multiplets():
    h = H9()
    for cb, ops in zip(h.bases, h.ops):
        for sec in cb.sectors():
            V = cb[sec]
            assert np.abs(V.T @ V - np.eye(V.shape[1])).max() < 1e-12
            S2 = ops["S2"][(sec, sec)]
            ev = np.diag(S2)
            assert np.abs(S2 - np.diag(ev)).max() < 1e-10
            s = (-1 + np.sqrt(1 + 4 * ev)) / 2
            assert np.abs(2 * s - np.round(2 * s)).max() < 1e-8
            assert np.all(s >= abs(sec[0] - sec[1]) / 2 - 1e-8)


def test_possible_spin_focksectors_brute_force():
    for ref in [((2, 1), (1, 2), (2, 1)), ((3, 0), (1, 1), (0, 2)), ((2, 2), (1, 0))]:
        got = possible_spin_focksectors(ref)
        Na, Nb = sum(f[0] for f in ref), sum(f[1] for f in ref)
        per = []
        for na, nb in ref:
            n, m = na + nb, abs(na - nb)
            per.append([(a, n - a) for a in range(n + 1) if abs(2 * a - n) <= m])
        want = [tuple(reversed(t)) for t in itertools.product(*reversed(per))
                if sum(x[0] for x in t) == Na and sum(x[1] for x in t) == Nb]
        assert sorted(got) == sorted(want) and got[0] == ref
        assert len(set(got)) == len(got)


@pytest.mark.slow
def test_h9_openshell_tpsci_pins(oracle_lib):
    h = H9()
    eng = OracleEngine()
    e0, v0 = tpsci_ci(h.start(), h.ops, h.H, engine=eng, verbose=0, **TPSCI_ARGS)
    assert np.abs(e0 - E0_REF).max() < 1e-8
    e2 = compute_pt2_energy(v0, h.ops, h.H, thresh_foi=PT2_THRESH, engine=eng, verbose=0)
    assert np.abs(e0 + e2 - E02_REF).max() < 1e-8
    e0s, v0s = tpsci_ci(h.start(), h.ops, h.H, engine=eng, verbose=0, thresh_spin=1e-5,
                        **TPSCI_ARGS)
    assert np.abs(e0s - E0_SPIN_REF).max() < 1e-8
    assert len(v0s) > len(v0)
    # the spin-completed states are S^2 eigenstates (two doublets, a quartet)
    s2 = compute_s2(v0s, h.ops, engine=eng)
    assert np.abs(s2 - np.array([0.75, 0.75, 3.75])).max() < 1e-8

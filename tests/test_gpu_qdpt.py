"""Plan-free diagonal (fcg_space_diagonal) and QDPT (compute_qdpt_energy)
on the GPU.

Diagonal: against the C restatement of compute_diagonal
(src/tpsci_outer.jl:694-715) on whole Fock blocks, on a sparse rank <= 1
space, with the H0 (Hcmf) operator, and on a space whose boxes a plan has
already formed; and against the plan's stored diagonal (fcg_diagonal).
Bar 1e-12 relative (FP64, the same terms summed in the same order per
configuration).

QDPT: the reference's Cr2 case (test/test_qdpt.jl:40, atol 1e-12) through
the product path (spin-adapted bases, tps_ci_direct and compute_qdpt_energy
on the GPU), and against the oracle's restatement (itself pinned to those
values and to the dense Loewdin partitioning) on a synthetic 4-cluster
case with 86 model configurations, screened and unscreened: 1e-11
relative, the same first-order space size.
"""
import numpy as np
import pytest

from fermicg_amd import _lib
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.operator import extract_1body_operator
from fermicg_amd.tpsci import (Engine, _full_space, compute_diagonal, compute_qdpt_energy,
                               first_order_focks, tps_ci_direct)
from oracle.pyoracle import Oracle, qdpt
from tests.pt2_cases import four_clusters, six_clusters_rank1, three_clusters_m4
from tests.qdpt_cases import QDPT_REF, THRESH_FOI, Cr2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.clear()


def _flat(case, ham=None):
    return (flatten_ops(case.ops, [len(c) for c in case.clusters]),
            flatten_ham(ham if ham is not None else case.H, 4))


CASES = {"full_blocks_3c_M4": lambda: three_clusters_m4(R=1, full_blocks=True),
         "subset_4c": lambda: four_clusters(seed=6, R=2),
         "rank1_6c": lambda: six_clusters_rank1(R=2)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_space_diagonal_vs_oracle(eng, name):
    case = CASES[name]()
    st = case.state
    print("case", name, "configurations", len(st), flush=True)
    fops, fham = _flat(case)
    ref = Oracle(fops, fham, st).diagonal()
    hd = compute_diagonal(st, case.ops, case.H, engine=eng)
    scale = np.abs(ref).max()
    assert np.abs(hd - ref).max() < 1e-12 * scale
    # the H0 (Hcmf) operator, as the PT1 denominators use it
    h0 = extract_1body_operator(case.H, op_string="Hcmf")
    ref0 = Oracle(fops, flatten_ham(h0, 4), st).diagonal()
    hd0 = compute_diagonal(st, case.ops, h0, engine=eng)
    assert np.abs(hd0 - ref0).max() < 1e-12 * np.abs(ref0).max()
    # a space whose boxes a plan formed; the plan's stored diagonal
    sp = _lib.DeviceSpace.tpsci(eng.ctx, st)
    pl = _lib.Plan.create(eng.ctx, eng.ham(case.H), eng.ops(case.ops), sp)
    boxed = _lib.space_diagonal(eng.ctx, eng.ham(case.H), eng.ops(case.ops), sp)
    assert np.abs(boxed - ref).max() < 1e-12 * scale
    assert np.abs(pl.diagonal() - ref).max() < 1e-12 * scale
    del pl


def test_space_diagonal_rejects_tucker_space(eng):
    from fermicg_amd.bst import bst_pq_state
    case = three_clusters_m4(R=1)
    focks = sorted({tuple(map(tuple, f)) for f in case.state.fock.tolist()})
    b = bst_pq_state(case.bases, focks, case.ref, R=1)
    sp = _lib.DeviceSpace.bst(eng.ctx, b.flat())
    with pytest.raises(_lib.FCGError):
        _lib.space_diagonal(eng.ctx, eng.ham(case.H), eng.ops(case.ops), sp)


def test_qdpt_cr2_reference_values(eng):
    p = Cr2()
    st = p.start()
    e0, v, _ = tps_ci_direct(st, p.ops, p.H, engine=eng)
    print("e0", e0, flush=True)
    e = compute_qdpt_energy(v, p.ops, p.H, nbody=4, H0="Hcmf", thresh_foi=THRESH_FOI, verbose=1,
                            engine=eng)
    print("dE vs test_qdpt.jl:40", e - QDPT_REF, compute_qdpt_energy.last_stats, flush=True)
    assert len(e) == 4
    assert np.abs(e - QDPT_REF).max() < 1e-12


@pytest.mark.parametrize("open_method", ["dense", "sparse"])
@pytest.mark.parametrize("thresh", [0.0, 1e-3])
def test_qdpt_vs_oracle(eng, thresh, open_method):
    case = four_clusters(seed=6, R=2)
    st = case.state
    fops, fham = _flat(case)
    target = _full_space(case.ops, first_order_focks(st, case.ops, case.H), st.R)
    ref, nx = qdpt(fops, fham, st, target, thresh)
    e = compute_qdpt_energy(st, case.ops, case.H, thresh_foi=thresh, verbose=0, engine=eng,
                            open_method=open_method)
    print("thresh", thresh, open_method, "model", len(st), "first-order", nx,
          compute_qdpt_energy.last_stats, flush=True)
    assert compute_qdpt_energy.last_stats["n_first_order"] == nx
    assert np.abs(e - ref).max() < 1e-11 * np.abs(ref).max()

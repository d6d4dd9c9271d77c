"""Apply variants: every variant the autotuner may pick gives the same sigma,
bit for bit, as the one it picked (the variants change the workgroup shape
and the loads in flight, never the order of a sum).

Cases: a full-block TPSCI space (multi-row boxes), a thinned one (one-row
boxes on the sparse plan), a space with both, and a rank-1 space of 39,313
rows (many one-row boxes per workgroup for every one-row variant),
at FP64, FP32, BF16X2 and BF16, R = 2 and R = 6 (above the tuned root count,
where only the default exists and pinning anything else is refused).  The
FP64 result of the small cases is also checked against the oracle at 1e-12
relative.
"""
import numpy as np
import pytest

from fermicg_amd import _lib
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import Engine, cluster_sizes_of
from fermicg_amd.workloads import build_problem, rank1_space, tpsci_small
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

PRECS = ("fp64", "fp32", "bf16x2", "bf16")


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


def _space(kind, R):
    if kind == "rank1":
        p = build_problem(24, 6, 16, 48, 7, max_transfer=2)
        st = rank1_space(p, R=R, frac=0.3, seed=7)
        st.set_vector(np.random.Generator(np.random.PCG64(1)).normal(size=(len(st), R)))
        return p, st
    p, st = tpsci_small(seed=0, R=R)
    if kind == "full":
        return p, st
    rng = np.random.Generator(np.random.PCG64(3))
    blocks = []
    for k, (fi, a, b) in enumerate(st.blocks()):
        if kind == "thinned" or k % 2:
            idx = [i for i in range(a, b) if rng.random() > 0.3] or [a]
            if b - a > 1 and len(idx) == b - a:
                idx = idx[1:]
        else:
            idx = list(range(a, b))
        blocks.append((st.focks()[fi], [tuple(int(x) for x in st.cfg[i]) for i in idx]))
    sp = TPSCIstate.from_blocks(3, blocks, R=R)
    sp.set_vector(rng.normal(size=(len(sp), R)))
    return p, sp


@pytest.mark.parametrize("kind", ["full", "thinned", "mixed", "rank1"])
@pytest.mark.parametrize("R", [2, 6])
def test_variants_identical(eng, kind, R):
    p, st = _space(kind, R)
    c = st.coeffs
    print("case", kind, "R", R, "rows", len(st), flush=True)
    ref = None if kind == "rank1" else Oracle(
        flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
        flatten_ham(p.clustered_ham), st).sigma(c)
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham, precisions=PRECS)
    for prec in PRECS:
        tuned = pl.sigma(c, precision=prec)
        vm, vr = pl.apply_variant(prec, R)
        print(prec, "tuned variant", (vm, vr), flush=True)
        assert vm >= 0 and vr >= 0
        if prec == "fp64" and ref is not None:
            assert np.abs(tuned - ref).max() < 1e-12 * np.abs(ref).max()
        nm, nr, max_r = _lib.apply_variant_counts()
        assert (nm, nr, max_r) == (5, 4, 4)
        pairs = [(m, r) for m in range(nm) for r in range(nr)] if R <= max_r else [(vm, vr)]
        for m, r in pairs:
            pl.set_apply_variant(prec, R, m, r)
            assert pl.apply_variant(prec, R) == (m, r)
            s = pl.sigma(c, precision=prec)
            assert np.array_equal(s, tuned), (prec, m, r)
        if R > 4:
            with pytest.raises(_lib.FCGError):
                pl.set_apply_variant(prec, R, (vm + 1) % 3, 0)
        with pytest.raises(_lib.FCGError):
            pl.set_apply_variant(prec, R, nm, 0)
        with pytest.raises(_lib.FCGError):
            pl.set_apply_variant(prec, R, 0, nr)

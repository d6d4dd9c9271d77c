"""Edge cases of the sigma plan and the solver through the C ABI.

Oracle: the C restatement of tps_ci_matvec / build_full_H (oracle/
fcg_oracle.c).  Bars: FP64 sigma 1e-12 relative (max-norm).  Cases:
  * the largest root count (16) in one launch, and 17 refused;
  * terms restricted to 1, 2 and 3 clusters (the reference's nbody keyword),
    on whole Fock blocks and on one-configuration rows;
  * a space of one configuration, and one whole Fock block alone;
  * errors: a precision the plan was not built with, a vector of the wrong
    length, a plan that does not fit the process budget (nothing allocated,
    footprint reported), linearly dependent Davidson start vectors.
"""
import numpy as np
import pytest

from fermicg_amd import _lib
from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import Engine, cluster_sizes_of
from fermicg_amd.workloads import tpsci_small
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


@pytest.fixture(scope="module")
def small():
    return tpsci_small(seed=0, R=1)


def _orc(p, st, nbody=4):
    return Oracle(flatten_ops(p.cluster_ops, cluster_sizes_of(p.cluster_ops)),
                  flatten_ham(p.clustered_ham, nbody), st)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _sub(st, pick, R, seed):
    """The configurations of `st` that pick(fock index, config index) keeps."""
    blocks = []
    focks = st.focks()
    for fi, a, b in st.blocks():
        confs = [tuple(int(x) for x in st.cfg[i]) for i in range(a, b) if pick(fi, i)]
        if confs:
            blocks.append((focks[fi], confs))
    sp = TPSCIstate.from_blocks(st.N, blocks, R=R)
    sp.set_vector(np.random.Generator(np.random.PCG64(seed)).normal(size=(len(sp), R)))
    return sp


def test_sixteen_roots(eng, small):
    p, st = small
    c = np.random.Generator(np.random.PCG64(3)).normal(size=(len(st), 16))
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham)
    s = pl.sigma(c)
    ref = _orc(p, st).sigma(c)
    print("R=16 rel", _rel(s, ref), flush=True)
    assert _rel(s, ref) < 1e-12
    with pytest.raises(_lib.FCGError):
        pl.sigma(np.zeros((len(st), 17)))


@pytest.mark.parametrize("nbody", [1, 2, 3])
@pytest.mark.parametrize("kind", ["full", "rows"])
def test_nbody(eng, small, nbody, kind):
    p, st = small
    if kind == "rows":
        # every Fock block thinned: one-configuration rows (the sparse plan)
        st = _sub(st, lambda fi, i: i % 3 != 1, R=2, seed=nbody)
    else:
        st.set_vector(np.random.Generator(np.random.PCG64(nbody)).normal(size=(len(st), 1)))
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham, nbody=nbody)
    s = pl.sigma(st.coeffs)
    ref = _orc(p, st, nbody).sigma(st.coeffs)
    print(kind, "nbody", nbody, "rel", _rel(s, ref), flush=True)
    assert _rel(s, ref) < 1e-12


def test_one_configuration_and_one_block(eng, small):
    p, st = small
    one = _sub(st, lambda fi, i: i == 0, R=2, seed=5)
    assert len(one) == 1
    s = eng.tpsci_plan(one, p.cluster_ops, p.clustered_ham).sigma(one.coeffs)
    H = _orc(p, one).full_H()
    assert H.shape == (1, 1)
    assert np.abs(s - H @ one.coeffs).max() <= 1e-12 * np.abs(H).max()
    fi0 = max(range(len(st.focks())), key=lambda f: st.fock_start[f + 1] - st.fock_start[f])
    blk = _sub(st, lambda fi, i: fi == fi0, R=1, seed=6)
    s = eng.tpsci_plan(blk, p.cluster_ops, p.clustered_ham).sigma(blk.coeffs)
    ref = _orc(p, blk).sigma(blk.coeffs)
    print("one block of", len(blk), "rel", _rel(s, ref), flush=True)
    assert _rel(s, ref) < 1e-12


def test_errors(eng, small):
    p, st = small
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham)  # FP64 only
    with pytest.raises(_lib.FCGError):
        pl.sigma(st.coeffs, precision="fp32")
    with pytest.raises(ValueError):
        pl.sigma(np.zeros((len(st) + 1, 1)))
    with pytest.raises(_lib.FCGError):
        pl.davidson(np.ones((len(st), 2)))  # two equal start vectors
    # a plan larger than the per-call cap: refused, nothing allocated
    ops, ham = eng.ops(p.cluster_ops), eng.ham(p.clustered_ham)
    sp = _lib.DeviceSpace.tpsci(eng.ctx, st)
    none, need = _lib.Plan.create_limited(eng.ctx, ham, ops, sp, 1024)
    assert none is None and need > 1024
    _, held0 = _lib.plan_budget()
    ok, need2 = _lib.Plan.create_limited(eng.ctx, ham, ops, sp, need)
    assert ok is not None and need2 == need
    # the footprint admitted includes the build-only pair and task tables;
    # once built, the plan holds exactly its reported device bytes
    assert 0 < ok.device_bytes() <= need
    assert _lib.plan_budget()[1] == held0 + ok.device_bytes()
    s = ok.sigma(st.coeffs)
    assert _rel(s, _orc(p, st).sigma(st.coeffs)) < 1e-12
    del ok
    assert _lib.plan_budget()[1] == held0

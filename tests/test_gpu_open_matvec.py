"""open_matvec_thread (src/tpsci_matvec_thread.jl:11-129) on the GPU through the
C ABI (a screened pair plan over the first-order Fock blocks).

Oracle: orc_tpsci_open_sigma (oracle/fcg_oracle.c), pinned at thresh 0 to
dense H c in tests/test_open_sigma_oracle.py.  Bars: the same target Fock
blocks in the same order, the same inserted configurations (keyed by Fock
block and configuration), and sigma within 1e-12 relative (max-norm), at
thresh 0, 1e-9, 1e-4 and 1e-2, for 4- and 2-body terms.
"""
import numpy as np
import pytest

from fermicg_amd.marshal import flatten_ham, flatten_ops
from fermicg_amd.tpsci import Engine, _full_space, first_order_focks, open_matvec_thread
from oracle.pyoracle import Oracle
from tests.pt2_cases import four_clusters, three_clusters_m4

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return Engine(0)


def _keyed(st):
    focks = st.focks()
    return {(focks[fi], tuple(int(x) for x in st.cfg[i])): i
            for fi, a, b in st.blocks() for i in range(a, b)}


@pytest.mark.parametrize("name,nbody,thresh", [("three", 4, 0.0), ("three", 4, 1e-9),
                                               ("three", 4, 1e-4), ("three", 4, 1e-2),
                                               ("four", 4, 0.0), ("four", 4, 1e-2),
                                               ("four", 2, 1e-4)])
def test_open_matvec(eng, name, nbody, thresh):
    case = three_clusters_m4() if name == "three" else four_clusters()
    st = case.state
    print("case", name, "nbody", nbody, "thresh", thresh, "n_var", len(st), flush=True)
    out = open_matvec_thread(st, case.ops, case.H, thresh=thresh, nbody=nbody, engine=eng)
    focks = first_order_focks(st, case.ops, case.H)
    assert out.focks() == focks
    target = _full_space(case.ops, focks, st.R)
    orc = Oracle(flatten_ops(case.ops, [len(c) for c in case.clusters]),
                 flatten_ham(case.H, nbody), st)
    sig, hit = orc.open_sigma(st.coeffs, focks, np.diff(target.fock_start), nbody=nbody,
                              thresh=thresh)
    ref_keys = [k for k, i in _keyed(target).items() if hit[i]]
    got = _keyed(out)
    print("inserted", len(got), "oracle", len(ref_keys), "of", len(target), flush=True)
    assert list(got) == ref_keys
    ref = sig[hit]
    assert out.coeffs.shape == ref.shape
    err = np.abs(out.coeffs - ref).max() / np.abs(ref).max()
    print("rel", err, flush=True)
    assert err < 1e-12

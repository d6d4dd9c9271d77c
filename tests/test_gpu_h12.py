"""The reference's h12 TPSCI values on the GPU (test/test_tpsci.jl "tpsci
h12 64bit", :88-133): tpsci_ci with the reference's arguments, every sigma,
open matvec, diagonal and solve in libfcg_hip.so, must reach the reference's
e0 and, with compute_pt2_energy at thresh_foi 1e-10, its e0 + e2, at the
reference's 1e-8.  The open matvec's inserted keys are compared with the
oracle's set exactly on these molecular (spin-symmetric) integrals, where
contributions can cancel."""
import numpy as np
import pytest

from fermicg_amd.tpsci import (Engine, _full_space, compute_pt2_energy, first_order_focks,
                               open_matvec_thread, tpsci_ci)
from tests.h12_cases import E0_REF, E02_REF, PT2_THRESH, TPSCI_ARGS, h12_problem, h12_start
from tests.oracle_engine import OracleEngine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h12():
    return h12_problem()


def test_gpu_h12_tpsci_and_pt2_match_reference(h12):
    p, _ = h12
    eng = Engine(0)
    e0, v0 = tpsci_ci(h12_start(p), p.ops, p.H, engine=eng, verbose=0, **TPSCI_ARGS)
    assert np.abs(e0 - E0_REF).max() < 1e-8
    e2 = compute_pt2_energy(v0, p.ops, p.H, thresh_foi=PT2_THRESH, engine=eng, verbose=0)
    assert np.abs(e0 + e2 - E02_REF).max() < 1e-8


@pytest.mark.parametrize("thresh", [0.0, 1e-5, 1e-3])
def test_gpu_open_matvec_keys_h12(h12, thresh, oracle_lib):
    p, _ = h12
    st = h12_start(p)
    # a dense 7-root state over the start configurations and their first
    # TPSCI expansion, so many terms meet on each target configuration
    rng = np.random.default_rng(7)
    st.set_vector(rng.normal(size=(len(st), 7)))
    eng = Engine(0)
    out = open_matvec_thread(st, p.ops, p.H, thresh=thresh, engine=eng)
    focks = first_order_focks(st, p.ops, p.H)
    target = _full_space(p.ops, focks, 7)
    sig, hit = OracleEngine().open_sigma(st, target, p.ops, p.H, thresh=thresh)
    keys_gpu = {(tuple(map(tuple, out.fock[fi])), tuple(out.cfg[i]))
                for fi, a, b in out.blocks() for i in range(a, b)}
    keys_orc = {(tuple(map(tuple, target.fock[fi])), tuple(target.cfg[i]))
                for fi, a, b in target.blocks() for i in range(a, b) if hit[i]}
    assert keys_gpu == keys_orc
    assert np.abs(out.coeffs - sig[hit]).max() <= 1e-12 * max(1.0, np.abs(sig).max())

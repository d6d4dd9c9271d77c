"""fcg_sigma_device_noalloc: the sigma of fcg_sigma_device with its
interleave scratch in a fixed device array instead of an allocation.

Through the C ABI, on device-resident operands: two calls in a row (the
second on a smaller plan and root count) must equal fcg_sigma_device's
results bit for bit in FP64, FP32 and FP32_BF16C (same kernels, same order
of every sum; FP32_BF16C's bf16 copy of c shares the scratch with the FP32
copy that multi-row boxes stage); R x n_var above the fixed 2^23 values is
refused with FCG_EOVERFLOW and a message, and a call within it on the same
large space still equals fcg_sigma_device.
"""
import numpy as np
import pytest
import torch

from fermicg_amd import _lib
from fermicg_amd.state import TPSCIstate
from fermicg_amd.tpsci import Engine
from fermicg_amd.workloads import build_problem, rank1_space, tpsci_small

pytestmark = pytest.mark.gpu


def _both(pl, c, prec, ctx=None):
    dev = torch.device("cuda:0")
    cd = torch.from_numpy(np.ascontiguousarray(c.T)).to(dev)
    s1 = torch.full_like(cd, np.nan)
    s2 = torch.full_like(cd, np.nan)
    R = c.shape[1]
    torch.cuda.synchronize()
    pl.sigma_device(cd.data_ptr(), s1.data_ptr(), R, prec)
    pl.ctx.synchronize()
    pl.sigma_device_noalloc(cd.data_ptr(), s2.data_ptr(), R, prec, ctx=ctx)
    (ctx or pl.ctx).synchronize()
    return s1.cpu().numpy(), s2.cpu().numpy()


def test_noalloc_equals_sigma_device_two_calls():
    eng = Engine(0)
    p, st = tpsci_small(seed=0, R=4)
    st.set_vector(np.random.default_rng(1).normal(size=(len(st), 4)))
    precs = ("fp64", "fp32", "fp32_bf16c")
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham, precisions=precs)
    for prec in precs:
        a, b = _both(pl, st.coeffs, prec)
        print("first", len(st), prec, flush=True)
        assert not np.isnan(b).any() and np.array_equal(a, b)
    # second call: a smaller space and fewer roots, on another context
    p2 = build_problem(8, 2, 4, 16, seed=5)
    st2 = TPSCIstate.full_blocks(p2.bases, p2.focks, R=2)
    st2.set_vector(np.random.default_rng(2).normal(size=(len(st2), 2)))
    pl2 = eng.tpsci_plan(st2, p2.cluster_ops, p2.clustered_ham, precisions=precs)
    other = _lib.Context(0)
    for prec in precs:
        a, b = _both(pl2, st2.coeffs, prec, ctx=other)
        print("second", len(st2), prec, flush=True)
        assert not np.isnan(b).any() and np.array_equal(a, b)
    del pl2
    eng.drop_plans()
    other.close()


def test_noalloc_limit():
    eng = Engine(0)
    # 580,673 rank-1 configurations; one-body terms only keep the plan small
    p = build_problem(32, 8, 16, 64, seed=4242, max_transfer=2)
    st = rank1_space(p, R=4, frac=1.0, seed=4242)
    st.set_vector(np.random.default_rng(3).normal(size=(len(st), 4)))
    pl = eng.tpsci_plan(st, p.cluster_ops, p.clustered_ham, nbody=1,
                        precisions=("fp64", "fp32_bf16c"))
    n = len(st)
    assert 16 * n > 2 ** 23 >= 4 * n
    dev = torch.device("cuda:0")
    big = torch.zeros((16, n), dtype=torch.float64, device=dev)
    with pytest.raises(_lib.FCGError, match="FCG_EOVERFLOW"):
        pl.sigma_device_noalloc(big.data_ptr(), big.data_ptr(), 16)
    for prec in ("fp64", "fp32_bf16c"):
        a, b = _both(pl, st.coeffs, prec)
        assert not np.isnan(b).any() and np.array_equal(a, b)
    del pl
    eng.drop_plans()

"""Keyed TPSCIstate algebra (the reference's OrderedDict operations,
src/type_TPSCIstate.jl and src/tpsci_outer.jl:869-889), on the CPU.

Each operation is checked against a dictionary built from the states' keys
(Fock block, ClusterConfig): add! sums by key and appends new keys after the
existing ones, clip! drops configurations whose every root is below the
threshold and then empty blocks, project_out! removes keys and the blocks it
empties, overlap sums c1[k] c2[k]^T over common keys.
"""
import numpy as np

from fermicg_amd.state import (TPSCIstate, add_states, clip_state, overlap_states,
                               project_out_state, scale_state, zero_state)


def _mk(blocks, seed, R=2):
    s = TPSCIstate.from_blocks(2, blocks, R=R)
    s.set_vector(np.random.Generator(np.random.PCG64(seed)).normal(size=(len(s), R)))
    return s


def _d(s):
    focks = s.focks()
    return {(focks[fi], tuple(int(x) for x in s.cfg[i])): s.coeffs[i]
            for fi, a, b in s.blocks() for i in range(a, b)}


F1, F2, F3 = ((1, 0), (0, 1)), ((0, 1), (1, 0)), ((1, 1), (0, 0))
A = _mk([(F1, [(1, 1), (2, 1), (1, 2)]), (F2, [(1, 1), (2, 2)])], 1)
B = _mk([(F2, [(2, 2), (3, 1)]), (F3, [(1, 1)]), (F1, [(1, 2)])], 2)


def test_add():
    s = add_states(A, B)
    da, db, ds = _d(A), _d(B), _d(s)
    assert set(ds) == set(da) | set(db)
    for k, v in ds.items():
        ref = da.get(k, 0.0) + db.get(k, 0.0)
        assert np.array_equal(v, ref)
    # A's keys first, in order; then B's new keys in their blocks, new block last
    assert s.focks() == [F1, F2, F3]
    assert [tuple(c) for c in s.cfg[:3]] == [(1, 1), (2, 1), (1, 2)]
    assert [tuple(c) for c in s.cfg[3:6]] == [(1, 1), (2, 2), (3, 1)]


def test_clip():
    s = A
    thr = np.sort(np.abs(s.coeffs).max(axis=1))[2]  # drops the two smallest rows
    c = clip_state(s, thr)
    kept = {k for k, v in _d(s).items() if not np.all(np.abs(v) < thr)}
    assert set(_d(c)) == kept and len(c) == len(s) - 2
    # a block emptied by clipping is removed
    z = zero_state(A)
    z.coeffs[:3] = 1.0
    assert clip_state(z, 0.5).focks() == [F1]


def test_project_out_and_overlap():
    p = project_out_state(A, B)
    assert set(_d(p)) == set(_d(A)) - set(_d(B))
    one = _mk([(F2, [(1, 1), (2, 2)])], 3)
    assert project_out_state(A, one).focks() == [F1]  # F2 emptied and dropped
    S = overlap_states(A, B)
    da, db = _d(A), _d(B)
    ref = sum(np.outer(da[k], db[k]) for k in set(da) & set(db))
    assert np.allclose(S, ref, rtol=0, atol=1e-15)
    M = np.array([[0.5, 1.0], [-2.0, 0.25]])
    assert np.allclose(scale_state(A, M).coeffs, A.coeffs @ M, rtol=0, atol=1e-15)

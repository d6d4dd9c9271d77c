"""Transition densities pinned by the reference's own values.

test/test_TDMs.jl:55-170 builds complete local bases for h6 (the decoded
_testdata_hf_h6.jld2 integrals) with clusters (1:4), (5:6) and asserts
seven slices |Γ[p, (q,) :, :]| of the A, B, AA, aa, Aa, Bb and Ab
operators to 1e-5 (absolute values: the eigenvectors' signs are
arbitrary).  The slices are kept as data in tests/golden/tdm_h6.json
(written from that file by tests/golden/extract_tdm_values.py); here the
same slices of fermicg_amd's ClusterOps producer (local FCI per sector,
cluster_ops.compute_cluster_ops) must match.  A leading index over several
orbitals is flattened column-major, p + n q, as the reference's arrays are
laid out.
"""
import json
import os

import numpy as np

from fermicg_amd.cluster_ops import compute_cluster_eigenbasis, compute_cluster_ops
from fermicg_amd.integrals import load_fixture

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_tdm_slices_match_reference():
    f = load_fixture(os.path.join(GOLD, "h6.npz"))
    clusters = [[0, 1, 2, 3], [4, 5]]
    bases = compute_cluster_eigenbasis(f["ints"], clusters, max_roots=100)
    ops = compute_cluster_ops(bases, f["ints"])
    recs = json.load(open(os.path.join(GOLD, "tdm_h6.json")))
    assert len(recs) == 7
    for r in recs:
        n = len(clusters[r["cluster"]])
        p = 0
        for k, i in enumerate(r["lead"]):
            p += i * n ** k
        G = ops[r["cluster"]][r["op"]][(tuple(r["bra"]), tuple(r["ket"]))]
        ref = np.array(r["ref"])
        got = G[p]
        print(r["op"], r["bra"], r["ket"], np.abs(np.abs(got) - np.abs(ref)).max())
        assert got.shape == ref.shape
        assert np.abs(np.abs(got) - np.abs(ref)).max() < 1e-5

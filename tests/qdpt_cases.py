"""The reference's Cr2 QDPT case (test/test_qdpt.jl): fixture integrals and
CMF 1-RDMs of test/_testdata_cmf_cr2_19.jld2 (tests/golden/cr2_19.npz;
19 orbitals in clusters of 5, 5, 3, 3, 3), spin-adapted cluster bases with
delta_elec 3 on every cluster and max_roots 20, CMF operators, four roots
in the spin-completed reference block (one all-ground configuration in each
of the four Fock blocks with 3 alpha + 3 beta electrons shared between the
first two clusters), solved by tps_ci_direct, and the QDPT energies that
test pins (test_qdpt.jl:40, atol 1e-12; energies without the nuclear
repulsion, which the ClusteredOperator leaves out)."""
import os

import numpy as np

from fermicg_amd.cluster_ops import (add_cmf_operators, compute_cluster_eigenbasis_spin,
                                     compute_cluster_ops)
from fermicg_amd.integrals import load_fixture
from fermicg_amd.operator import extract_ClusteredTerms
from fermicg_amd.state import TPSCIstate, add_spin_focksectors

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# test/test_qdpt.jl:40
QDPT_REF = np.array([-108.13989647019216, -108.13975457069576, -108.1394893503325,
                     -108.1390882840833])
THRESH_FOI = 1e-8   # test_qdpt.jl:34
M = 20              # test_qdpt.jl:12


class Cr2:
    def __init__(self):
        f = load_fixture(os.path.join(GOLD, "cr2_19.npz"))
        self.ints = f["ints"]
        self.clusters = f["clusters"]
        self.ref = tuple(tuple(int(y) for y in x) for x in f["init_fspace"])
        self.bases = compute_cluster_eigenbasis_spin(self.ints, self.clusters, f["d1a"], f["d1b"],
                                                     [3] * 5, self.ref, max_roots=M)
        self.ops = compute_cluster_ops(self.bases, self.ints)
        add_cmf_operators(self.ops, self.bases, self.ints, f["d1a"], f["d1b"])
        self.H = extract_ClusteredTerms(self.ints, self.clusters)

    def start(self, R=4):
        st = TPSCIstate.from_blocks(5, [(self.ref, [(1,) * 5])], R=R)
        return add_spin_focksectors(st)

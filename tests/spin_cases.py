"""The reference's h9 open-shell TPSCI case (test/test_openshell.jl
"openshell_tpsci", :9-61): fixture integrals and CMF 1-RDMs of
test/_testdata_cmf_h9.jld2 (tests/golden/h9.npz), spin-adapted cluster
bases with delta_elec [3,3,3] and max_roots 100, three roots started from
the spin-completed reference block, and the energies that test pins."""
import os

import numpy as np

from fermicg_amd.cluster_ops import (add_cmf_operators, compute_cluster_eigenbasis_spin,
                                     compute_cluster_ops)
from fermicg_amd.integrals import load_fixture
from fermicg_amd.operator import extract_ClusteredTerms
from fermicg_amd.state import TPSCIstate, add_spin_focksectors

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# test/test_openshell.jl:38-44 (e0, 8 decimals), :38-44 e2_ref (e0 + e2)
E0_REF = np.array([-14.05014150, -14.02155292, -14.00595447])
E02_REF = np.array([-14.05028658, -14.02164792, -14.00602933])
# test/test_openshell.jl:52-61: the same with thresh_spin = 1e-5
E0_SPIN_REF = np.array([-14.050151820221835, -14.021560215779372, -14.005960750225363])
TPSCI_ARGS = dict(incremental=True, thresh_cipsi=1e-3, thresh_foi=1e-5, thresh_asci=-1)
PT2_THRESH = 1e-8


class H9:
    def __init__(self):
        f = load_fixture(os.path.join(GOLD, "h9.npz"))
        self.ints = f["ints"]
        self.clusters = f["clusters"]
        self.ref = tuple(tuple(x) for x in f["init_fspace"])
        self.bases = compute_cluster_eigenbasis_spin(self.ints, self.clusters, f["d1a"], f["d1b"],
                                                     [3, 3, 3], self.ref, max_roots=100)
        self.ops = compute_cluster_ops(self.bases, self.ints)
        add_cmf_operators(self.ops, self.bases, self.ints, f["d1a"], f["d1b"])
        self.H = extract_ClusteredTerms(self.ints, self.clusters)

    def start(self, R=3):
        st = TPSCIstate.from_blocks(3, [(self.ref, [(1, 1, 1)])], R=R)
        return add_spin_focksectors(st)

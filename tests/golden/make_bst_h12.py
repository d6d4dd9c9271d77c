"""Writes tests/golden/bst_h12_oracle.json: block_sparse_tucker on the
reference's h12 BST case (test/test_bst.jl:72-88 arguments) run on the CPU
oracle engine (the C restatement of build_sigma! for every Hamiltonian
application, dense eigensolves), for the GPU test to compare with.
Takes several minutes on 8 threads.  Usage: python tests/golden/make_bst_h12.py"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from fermicg_amd.bst_loop import block_sparse_tucker, pq_spaces, reference_state  # noqa: E402
from tests.bst_loop_cases import H12_ARGS  # noqa: E402
from tests.h12_cases import h12_problem  # noqa: E402
from tests.oracle_engine import OracleEngine  # noqa: E402


def main():
    p, _ = h12_problem()
    e, _ = block_sparse_tucker(reference_state(5, p.ref), p.ops, p.H, pq_spaces(p.bases, p.ref),
                               engine=OracleEngine(nthreads=8), verbose=1, **H12_ARGS)
    hist = block_sparse_tucker.history
    out = {"e_var": float(e[0]),
           "dims": [[h["dim_ref"], h["dim_foi"], h["dim_foi_compressed"], h["dim_var"]]
                    for h in hist],
           "e_ref": [float(h["e_ref"][0]) for h in hist],
           "e_pt2": [float(h["e_pt2"][0]) for h in hist],
           "e_iter": [float(h["e_var"][0]) for h in hist]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bst_h12_oracle.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Decode the reference's JLD2 test fixtures into plain .npz data files.

JLD2 files are HDF5 containers; this reads them with h5py (data only,
nothing from the file is executed) and writes one ``<name>.npz`` per fixture
next to this script.  Julia stores arrays column-major, so every array read
through h5py comes back with its axes reversed and is transposed here.

Run with an interpreter that has h5py (not needed at test time; the .npz
files are committed):

    /opt/conda/bin/python3.9 tests/golden/decode_jld2.py /root/reference/test

Fixtures decoded (reference test/ directory):
  _testdata_hf_h6.jld2       ints, clusters, init_fspace, e_fci
  _testdata_cmf_h8.jld2      ints, d1, clusters, init_fspace, e_fci
  _testdata_cmf_h9.jld2      ints, d1, clusters, init_fspace, e_fci (3 states)
  _testdata_cmf_h12_64bit.jld2  ints, d1, clusters, init_fspace, e_cmf
  _testdata_cmf_cr2_19.jld2  ints, d1, clusters, init_fspace (test/test_qdpt.jl)
The ``cluster_bases`` entries hold ActiveSpaceSolvers objects whose
determinant ordering is not documented in the reference; they are skipped.
"""
import os
import sys

import h5py
import numpy as np


def _arr(f, ref):
    a = f[ref][()]
    return np.ascontiguousarray(np.transpose(a))


def decode(path, out):
    f = h5py.File(path, "r")
    rec = {}
    ints = f["ints"][()]
    rec["h0"] = np.array(float(ints["h0"]))
    rec["h1"] = _arr(f, ints["h1"])
    rec["h2"] = _arr(f, ints["h2"])
    orb_lists = []
    for cref in f["clusters"][()]:
        c = f[cref][()]
        orbs = np.asarray(f[c[1]][()], dtype=np.int64).ravel()
        orb_lists.append(orbs)
    rec["cluster_sizes"] = np.array([len(o) for o in orb_lists], dtype=np.int64)
    rec["cluster_orbs"] = np.concatenate(orb_lists)  # 1-based, as in Julia
    fs = f["init_fspace"][()]
    rec["init_fspace"] = np.array([[int(x[0]), int(x[1])] for x in fs], dtype=np.int64)
    for key in ("e_fci", "e_cmf"):
        if key in f:
            rec[key] = np.atleast_1d(np.asarray(f[key][()], dtype=np.float64))
    if "d1" in f:
        d1 = f["d1"][()]
        rec["d1a"] = _arr(f, d1["a"])
        rec["d1b"] = _arr(f, d1["b"])
    np.savez(out, **rec)
    return rec


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test"
    here = os.path.dirname(os.path.abspath(__file__))
    names = {
        "_testdata_hf_h6.jld2": "h6",
        "_testdata_cmf_h8.jld2": "h8",
        "_testdata_cmf_h9.jld2": "h9",
        "_testdata_cmf_h12_64bit.jld2": "h12",
        "_testdata_cmf_cr2_19.jld2": "cr2_19",
    }
    for fn, tag in names.items():
        rec = decode(os.path.join(src, fn), os.path.join(here, tag + ".npz"))
        print(tag, {k: (v.shape if v.ndim else float(v)) for k, v in rec.items()})


if __name__ == "__main__":
    main()

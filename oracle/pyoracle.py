"""ctypes front end of the C oracle (test infrastructure only).

Loads oracle/liboracle.so (built by `make -C oracle`) and exposes the
reference-restated TPSCI sigma, build_full_H and diagonal on the flat wire
format of fermicg_amd.marshal.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

P = ctypes.POINTER
i32, i64, i16, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_int16, ctypes.c_double


class OrcOps(ctypes.Structure):
    _fields_ = [("n_clusters", ctypes.c_int), ("S", ctypes.c_int), ("n_ops", ctypes.c_int),
                ("gblock", P(i32)), ("gb_off", P(i64)), ("gb_dims", P(i32)), ("gdata", P(f64))]


class OrcHam(ctypes.Structure):
    _fields_ = [("n_terms", ctypes.c_int), ("nbody", P(i32)), ("clusters", P(i32)),
                ("ops", P(i32)), ("parity", P(i32)), ("dims", P(i32)), ("delta", P(i16)),
                ("ioff", P(i64)), ("ints", P(f64)), ("n_tc", ctypes.c_int), ("tc", P(i16)),
                ("tc_start", P(i32))]


class OrcSpace(ctypes.Structure):
    _fields_ = [("n_fock", i64), ("fock", P(i16)), ("fock_start", P(i64)), ("n_cfg", i64),
                ("cfg", P(i16))]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        L = ctypes.CDLL(path)
        for name in ("orc_tpsci_sigma",):
            getattr(L, name).restype = ctypes.c_int
        L.orc_tpsci_sigma.argtypes = [P(OrcOps), P(OrcHam), P(OrcSpace), ctypes.c_int, P(f64), P(f64), ctypes.c_int]
        L.orc_build_full_H.argtypes = [P(OrcOps), P(OrcHam), P(OrcSpace), P(f64), ctypes.c_int]
        L.orc_tpsci_diagonal.argtypes = [P(OrcOps), P(OrcHam), P(OrcSpace), ctypes.c_int, P(f64)]
        L.orc_tpsci_sigma_rows.argtypes = [P(OrcOps), P(OrcHam), P(OrcSpace), ctypes.c_int, P(f64),
                                           P(i64), i64, P(f64), ctypes.c_int]
        L.orc_tpsci_pt2.argtypes = [P(OrcOps), P(OrcHam), P(OrcSpace), ctypes.c_int, P(f64), P(f64),
                                    ctypes.c_int, ctypes.c_int, f64, P(f64), P(i64), ctypes.c_int]
        L.orc_tpsci_pt2_sparse.argtypes = [P(OrcOps), P(OrcHam), P(OrcSpace), ctypes.c_int,
                                           P(f64), P(f64), ctypes.c_int, ctypes.c_int, f64,
                                           P(f64), P(i64), P(i64), P(i64), ctypes.c_int]
        L.orc_tpsci_open_sigma.argtypes = [P(OrcOps), P(OrcHam), P(OrcSpace), ctypes.c_int, P(f64),
                                           ctypes.c_int, f64, i64, P(ctypes.c_int16), P(i64), P(f64),
                                           P(ctypes.c_uint8), ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(P(t))


class Oracle:
    """Holds the ctypes structs (and keeps the numpy arrays alive)."""

    def __init__(self, fops, fham, state):
        self._keep = []

        def k(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        self.ops = OrcOps(fops.n_clusters, fops.S, fops.n_ops,
                          _p(k(fops.gblock, np.int32), i32), _p(k(fops.gb_off, np.int64), i64),
                          _p(k(fops.gb_dims, np.int32), i32), _p(k(fops.gdata, np.float64), f64))
        self.ham = OrcHam(fham.n_terms, _p(k(fham.nbody, np.int32), i32),
                          _p(k(fham.clusters, np.int32), i32), _p(k(fham.ops, np.int32), i32),
                          _p(k(fham.parity, np.int32), i32), _p(k(fham.dims, np.int32), i32),
                          _p(k(fham.delta, np.int16), i16), _p(k(fham.ioff, np.int64), i64),
                          _p(k(fham.ints, np.float64), f64), fham.n_tc,
                          _p(k(fham.tc, np.int16), i16), _p(k(fham.tc_start, np.int32), i32))
        self.n_cfg = len(state)
        self.space = OrcSpace(state.n_fock, _p(k(state.fock, np.int16), i16),
                              _p(k(state.fock_start, np.int64), i64), len(state),
                              _p(k(state.cfg, np.int16), i16))
        zero = tuple((0, 0) for _ in range(fops.n_clusters))
        self.zero_tc = -1
        for g in range(fham.n_tc):
            if all(tuple(x) == (0, 0) for x in fham.tc[g]):
                self.zero_tc = g

    def sigma(self, c, nthreads=0):
        c = np.asarray(c, dtype=np.float64)
        if c.ndim == 1:
            c = c[:, None]
        R = c.shape[1]
        cc = np.ascontiguousarray(c.T)
        out = np.zeros((R, self.n_cfg))
        rc = lib().orc_tpsci_sigma(ctypes.byref(self.ops), ctypes.byref(self.ham),
                                   ctypes.byref(self.space), R, _p(cc, f64), _p(out, f64), nthreads)
        if rc:
            raise RuntimeError("oracle sigma failed: %d" % rc)
        return out.T.copy()

    def sigma_rows(self, c, rows, nthreads=0):
        c = np.asarray(c, dtype=np.float64)
        if c.ndim == 1:
            c = c[:, None]
        R = c.shape[1]
        cc = np.ascontiguousarray(c.T)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.zeros((R, len(rows)))
        rc = lib().orc_tpsci_sigma_rows(ctypes.byref(self.ops), ctypes.byref(self.ham),
                                        ctypes.byref(self.space), R, _p(cc, f64), _p(rows, i64),
                                        len(rows), _p(out, f64), nthreads)
        if rc:
            raise RuntimeError("oracle sigma failed: %d" % rc)
        return out.T.copy()

    def full_H(self, nthreads=0):
        n = self.n_cfg
        H = np.zeros((n, n))
        rc = lib().orc_build_full_H(ctypes.byref(self.ops), ctypes.byref(self.ham),
                                    ctypes.byref(self.space), _p(H, f64), nthreads)
        if rc:
            raise RuntimeError("oracle build_full_H failed: %d" % rc)
        return H

    def pt2(self, c, e0, nbody=4, h0_op=1, thresh=0.0, nthreads=0):
        """compute_pt2_energy restated (prescreen=false); returns (e2, n_jobs)."""
        c = np.asarray(c, dtype=np.float64)
        if c.ndim == 1:
            c = c[:, None]
        R = c.shape[1]
        cc = np.ascontiguousarray(c.T)
        e0 = np.ascontiguousarray(np.broadcast_to(np.asarray(e0, np.float64), (R,)))
        e2 = np.zeros(R)
        nj = i64()
        rc = lib().orc_tpsci_pt2(ctypes.byref(self.ops), ctypes.byref(self.ham),
                                 ctypes.byref(self.space), R, _p(cc, f64), _p(e0, f64), nbody,
                                 h0_op, thresh, _p(e2, f64), ctypes.byref(nj), nthreads)
        if rc:
            raise RuntimeError("oracle pt2 failed: %d" % rc)
        return e2, nj.value

    def pt2_sparse(self, c, e0, nbody=4, h0_op=1, thresh=0.0, nthreads=0):
        """compute_pt2_energy restated over the configurations reached (target
        blocks never formed whole; zero-coefficient kets skipped); returns
        (e2, n_jobs, n_first_order); the kept elements are in
        self.last_contributions."""
        c = np.asarray(c, dtype=np.float64)
        if c.ndim == 1:
            c = c[:, None]
        R = c.shape[1]
        cc = np.ascontiguousarray(c.T)
        e0 = np.ascontiguousarray(np.broadcast_to(np.asarray(e0, np.float64), (R,)))
        e2 = np.zeros(R)
        nj, nx, nc = i64(), i64(), i64()
        rc = lib().orc_tpsci_pt2_sparse(ctypes.byref(self.ops), ctypes.byref(self.ham),
                                        ctypes.byref(self.space), R, _p(cc, f64), _p(e0, f64),
                                        nbody, h0_op, thresh, _p(e2, f64), ctypes.byref(nj),
                                        ctypes.byref(nx), ctypes.byref(nc), nthreads)
        if rc:
            raise RuntimeError("oracle pt2_sparse failed: %d" % rc)
        self.last_contributions = nc.value
        return e2, nj.value, nx.value

    def open_sigma(self, c, target_focks, sizes, nbody=4, thresh=0.0, nthreads=0):
        """open_matvec_thread restated (prescreen=false) on the listed target
        Fock blocks of `sizes` configurations each (first cluster fastest):
        returns (sigma [n x R], hit [n] bool: configurations the reference
        inserts)."""
        c = np.asarray(c, dtype=np.float64)
        if c.ndim == 1:
            c = c[:, None]
        R = c.shape[1]
        cc = np.ascontiguousarray(c.T)
        tf = np.ascontiguousarray(np.asarray(target_focks, np.int16).reshape(len(target_focks), -1))
        off = np.zeros(len(sizes) + 1, np.int64)
        off[1:] = np.cumsum(sizes)
        sig = np.zeros((R, int(off[-1])))
        hit = np.zeros(int(off[-1]), np.uint8)
        rc = lib().orc_tpsci_open_sigma(ctypes.byref(self.ops), ctypes.byref(self.ham),
                                        ctypes.byref(self.space), R, _p(cc, f64), nbody, thresh,
                                        len(sizes), _p(tf, ctypes.c_int16), _p(off, i64),
                                        _p(sig, f64), _p(hit, ctypes.c_uint8), nthreads)
        if rc:
            raise RuntimeError("oracle open_sigma failed: %d" % rc)
        return sig.T.copy(), hit.astype(bool)

    def diagonal(self):
        hd = np.zeros(self.n_cfg)
        rc = lib().orc_tpsci_diagonal(ctypes.byref(self.ops), ctypes.byref(self.ham),
                                      ctypes.byref(self.space), self.zero_tc, _p(hd, f64))
        if rc:
            raise RuntimeError("oracle diagonal failed: %d" % rc)
        return hd


class OrcBst(ctypes.Structure):
    _fields_ = [("n_fock", i64), ("fock", P(i16)), ("n_box", i64), ("box_fock", P(i32)),
                ("lo", P(i16)), ("hi", P(i16)), ("rank", P(i16)), ("foff", P(i64)),
                ("factors", P(f64)), ("var_off", P(i64))]


class BstOracle(Oracle):
    """Reference build_sigma! restated in C, on a fermicg_amd.bst.BSTstate."""

    def __init__(self, fops, fham, bst):
        self._keep = []

        def k(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        self.ops = OrcOps(fops.n_clusters, fops.S, fops.n_ops,
                          _p(k(fops.gblock, np.int32), i32), _p(k(fops.gb_off, np.int64), i64),
                          _p(k(fops.gb_dims, np.int32), i32), _p(k(fops.gdata, np.float64), f64))
        self.ham = OrcHam(fham.n_terms, _p(k(fham.nbody, np.int32), i32),
                          _p(k(fham.clusters, np.int32), i32), _p(k(fham.ops, np.int32), i32),
                          _p(k(fham.parity, np.int32), i32), _p(k(fham.dims, np.int32), i32),
                          _p(k(fham.delta, np.int16), i16), _p(k(fham.ioff, np.int64), i64),
                          _p(k(fham.ints, np.float64), f64), fham.n_tc,
                          _p(k(fham.tc, np.int16), i16), _p(k(fham.tc_start, np.int32), i32))
        fl = bst.flat()
        self.var_off = k(fl.var_off, np.int64)
        self.nvar = int(self.var_off[-1])
        self.n_box = fl.n_box
        self.bst = OrcBst(fl.n_fock, _p(k(fl.fock, np.int16), i16), fl.n_box,
                          _p(k(fl.box_fock, np.int32), i32), _p(k(fl.lo, np.int16), i16),
                          _p(k(fl.hi, np.int16), i16), _p(k(fl.rank, np.int16), i16),
                          _p(k(fl.foff, np.int64), i64),
                          _p(k(fl.factors if fl.factors.size else np.zeros(1), np.float64), f64),
                          _p(self.var_off, i64))
        L = lib()
        L.orc_bst_sigma_rows.argtypes = [P(OrcOps), P(OrcHam), P(OrcBst), ctypes.c_int, P(f64),
                                         i64, P(i64), i64, P(f64), ctypes.c_int, P(f64), P(f64)]

    def sigma_rows(self, c, rows, nthreads=0, timings=False):
        """sigma restricted to the bra blocks `rows`: returns (values, index
        array of the variables they cover[, (t_build, t_contract)])."""
        c = np.asarray(c, dtype=np.float64)
        if c.ndim == 1:
            c = c[:, None]
        R = c.shape[1]
        cc = np.ascontiguousarray(c.T)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        idx = np.concatenate([np.arange(self.var_off[b], self.var_off[b + 1]) for b in rows]) \
            if len(rows) else np.zeros(0, np.int64)
        out = np.zeros((R, len(idx)))
        tb, tc = f64(), f64()
        rc = lib().orc_bst_sigma_rows(ctypes.byref(self.ops), ctypes.byref(self.ham),
                                      ctypes.byref(self.bst), R, _p(cc, f64), self.nvar,
                                      _p(rows, i64), len(rows), _p(out, f64), nthreads,
                                      ctypes.byref(tb), ctypes.byref(tc))
        if rc:
            raise RuntimeError("oracle bst sigma failed: %d" % rc)
        if timings:
            return out.T.copy(), idx, (tb.value, tc.value)
        return out.T.copy(), idx

    def sigma(self, c, nthreads=0):
        s, idx = self.sigma_rows(c, np.arange(self.n_box), nthreads)
        return s


class _Configs:
    """Configurations of some Fock blocks in the TPSCIstate layout (fock,
    fock_start, cfg), for Oracle()."""

    def __init__(self, fock, fock_start, cfg):
        self.fock = np.ascontiguousarray(fock, np.int16)
        self.fock_start = np.ascontiguousarray(fock_start, np.int64)
        self.cfg = np.ascontiguousarray(cfg, np.int16)
        self.n_fock = self.fock.shape[0]

    def __len__(self):
        return self.cfg.shape[0]


def qdpt(fops, fham, state, target, thresh_foi, E0=None, nthreads=0):
    """compute_qdpt_energy restated (src/tpsci_pt2_energy.jl:311-361) for an
    orthonormal state, on the C restatements above:
      E0[r] = <r|H|r> unless given (:330-333);
      X = configurations of `target` (every configuration of the open
          matvec's target Fock blocks, first cluster fastest) that the
          screened open matvec at thresh_foi reaches (open_matvec_thread,
          :335-339), minus the state's own (project_out!, :340);
      Hd = diagonal of H over X (compute_diagonal, :344);
      H_ax = <a|H|x> for model configurations a (build_full_H_parallel,
          :345), here the open matvec of each unit vector without screening;
      H2 = H_ax pinv(E0[1] I - diag(Hd)) H_ax^T (:349-352), exact zeros of
          the diagonal left out as pinv does;
      eigenvalues of H2 + H_model (:354-357), ascending.
    Returns (energies, number of configurations in X)."""
    c = np.asarray(state.coeffs, np.float64)
    orc = Oracle(fops, fham, state)
    h_model = orc.full_H(nthreads)
    if E0 is None:
        E0 = np.einsum("ir,ij,jr->r", c, h_model, c)
    E0 = np.asarray(E0, np.float64).reshape(-1)
    focks = target.fock.reshape(target.n_fock, -1)
    sizes = np.diff(target.fock_start)
    _, hit = orc.open_sigma(c, focks, sizes, thresh=thresh_foi, nthreads=nthreads)
    model = {(state.fock[b].tobytes(), state.cfg[i].tobytes())
             for b in range(state.n_fock) for i in range(state.fock_start[b],
                                                         state.fock_start[b + 1])}
    blk = np.repeat(np.arange(target.n_fock), sizes)
    keep = hit.copy()
    for i in np.nonzero(hit)[0]:
        if (target.fock[blk[i]].tobytes(), target.cfg[i].tobytes()) in model:
            keep[i] = False
    xi = np.nonzero(keep)[0]
    if len(xi) == 0:
        return np.linalg.eigvalsh(h_model), 0
    starts = np.searchsorted(xi, target.fock_start)
    xs = _Configs(target.fock, starts, target.cfg[xi])
    hd = Oracle(fops, fham, xs).diagonal()
    n_model = c.shape[0]
    h_xa = np.zeros((len(xi), n_model))
    for j0 in range(0, n_model, 64):   # the restatement takes at most 64 vectors
        j1 = min(n_model, j0 + 64)
        unit = np.zeros((n_model, j1 - j0))
        unit[np.arange(j0, j1), np.arange(j1 - j0)] = 1.0
        h_all, _ = orc.open_sigma(unit, focks, sizes, thresh=0.0, nthreads=nthreads)
        h_xa[:, j0:j1] = h_all[xi]
    den = E0[0] - hd
    dinv = np.zeros_like(den)
    nz = den != 0.0
    dinv[nz] = 1.0 / den[nz]
    h2 = (h_xa.T * dinv[None, :]) @ h_xa
    h = h2 + h_model
    return np.linalg.eigvalsh(0.5 * (h + h.T)), len(xi)

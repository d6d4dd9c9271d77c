"""ctypes binding of libfcg_hip.so (the C ABI declared in include/fcg.h).

The product path always runs through this library; if it is missing or no
GPU is visible, every entry point raises -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfcg_hip.so")

FCG_FP64, FCG_FP32, FCG_BF16X2, FCG_BF16, FCG_FP32_BF16C = 0, 1, 2, 3, 4
# "fp32_bf16c": the FP32 blocks applied to c rounded to bf16
PRECISION = {"fp64": FCG_FP64, "fp32": FCG_FP32, "bf16x2": FCG_BF16X2, "bf16": FCG_BF16,
             "fp32_bf16c": FCG_FP32_BF16C}
ERRORS = {-1: "FCG_EINVAL", -2: "FCG_ENOMEM", -3: "FCG_EMISSING_BLOCK", -4: "FCG_EOVERFLOW",
          -5: "FCG_EDEVICE", -6: "FCG_ENOTCONV"}

P = ctypes.POINTER
c_i16, c_i32, c_i64, c_f64 = ctypes.c_int16, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
VP = ctypes.c_void_p

EXPORTS = [
    "fcg_create", "fcg_destroy", "fcg_last_error", "fcg_ops_upload", "fcg_ops_destroy",
    "fcg_ham_upload", "fcg_ham_destroy", "fcg_tpsci_space_upload", "fcg_bst_space_upload",
    "fcg_space_nvar", "fcg_space_destroy", "fcg_plan_create", "fcg_plan_destroy",
    "fcg_plan_stats", "fcg_sigma", "fcg_sigma_device", "fcg_diagonal", "fcg_davidson",
    "fcg_last_sigma_time", "fcg_event_record", "fcg_event_elapsed", "fcg_synchronize",
    "fcg_plan_create_pair", "fcg_tpsci_pt2", "fcg_plan_create_limited", "fcg_plan_device_bytes",
    "fcg_set_plan_budget", "fcg_plan_budget", "fcg_plan_create_pair_screened",
    "fcg_plan_apply_variant",
    "fcg_plan_set_apply_variant",
    "fcg_apply_variant_counts",
    "fcg_set_build_class_cap",
    "fcg_open_sigma",
    "fcg_open_sigma_sparse", "fcg_open_result_size", "fcg_open_result_copy",
    "fcg_open_result_destroy",
    "fcg_sigma_device_noalloc",
    "fcg_tpsci_pt2_sparse",
    "fcg_tpsci_pt2_sparse_ex",
    "fcg_plan_create_matrix_free",
    "fcg_serve_start", "fcg_serve_submit", "fcg_serve_poll", "fcg_serve_stop",
    "fcg_space_diagonal",
]


class DavidsonOpts(ctypes.Structure):
    _fields_ = [("conv_thresh", c_f64), ("max_iter", ctypes.c_int), ("max_ss_vecs", ctypes.c_int),
                ("lindep_thresh", c_f64), ("precision", ctypes.c_int), ("refresh_k", ctypes.c_int),
                ("precond", ctypes.c_int), ("precond_diag", P(c_f64))]


class DavidsonStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("sigma_fp64", ctypes.c_int),
                ("sigma_low", ctypes.c_int), ("converged", ctypes.c_int),
                ("residual", c_f64 * 16)]


class ServeOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int), ("hbm_budget", c_i64),
                ("nbody", ctypes.c_int), ("davidson", DavidsonOpts)]


class ServeTimes(ctypes.Structure):
    _fields_ = [("submitted", c_f64), ("started", c_f64), ("planned", c_f64),
                ("finished", c_f64), ("plan_bytes", c_i64)]


class PT2Stats(ctypes.Structure):
    _fields_ = [("n_target_fock", c_i64), ("n_target_var", c_i64), ("n_first_order", c_i64),
                ("block_pairs", c_i64), ("block_elems", c_i64), ("seconds", c_f64),
                ("build_flops", c_f64), ("build_seconds", c_f64)]


class PT2SparseStats(ctypes.Structure):
    _fields_ = [("n_items", c_i64), ("n_groups", c_i64), ("n_contributions", c_i64),
                ("n_configs", c_i64), ("n_first_order", c_i64), ("chain_flops", c_f64),
                ("seconds", c_f64), ("n_items_kept", c_i64), ("n_kets_kept", c_i64),
                ("n_chunks", c_i64), ("contraction_seconds", c_f64), ("sort_seconds", c_f64)]


class FCGError(RuntimeError):
    pass


_lib = None


def load():
    """Load the library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FCGError("libfcg_hip.so not built: run `make -C fermicg_amd/csrc` "
                       "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.fcg_create.argtypes = [ctypes.c_int, P(VP)]
    L.fcg_destroy.argtypes = [VP]
    L.fcg_destroy.restype = None
    L.fcg_last_error.argtypes = [VP]
    L.fcg_last_error.restype = ctypes.c_char_p
    L.fcg_ops_upload.argtypes = [VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(c_i32), c_i64,
                                 P(c_i64), P(c_i32), P(c_f64), c_i64, P(VP)]
    L.fcg_ops_destroy.argtypes = [VP]
    L.fcg_ops_destroy.restype = None
    L.fcg_ham_upload.argtypes = [VP, ctypes.c_int, ctypes.c_int, P(c_i32), P(c_i32), P(c_i32),
                                 P(c_i32), P(c_i32), P(c_i16), P(c_i64), P(c_f64), c_i64,
                                 ctypes.c_int, P(c_i16), P(c_i32), P(VP)]
    L.fcg_ham_destroy.argtypes = [VP]
    L.fcg_ham_destroy.restype = None
    L.fcg_tpsci_space_upload.argtypes = [VP, ctypes.c_int, c_i64, P(c_i16), P(c_i64), c_i64,
                                         P(c_i16), P(VP)]
    L.fcg_bst_space_upload.argtypes = [VP, ctypes.c_int, c_i64, P(c_i16), c_i64, P(c_i32),
                                       P(c_i16), P(c_i16), P(c_i16), P(c_i64), P(c_f64), c_i64,
                                       P(VP)]
    L.fcg_space_nvar.argtypes = [VP]
    L.fcg_space_nvar.restype = c_i64
    L.fcg_space_destroy.argtypes = [VP]
    L.fcg_space_destroy.restype = None
    L.fcg_plan_create.argtypes = [VP, VP, VP, VP, ctypes.c_int, ctypes.c_int, P(VP)]
    L.fcg_plan_create_pair.argtypes = [VP, VP, VP, VP, VP, ctypes.c_int, ctypes.c_int, P(VP)]
    L.fcg_tpsci_pt2.argtypes = [VP, VP, VP, VP, ctypes.c_int, P(c_f64), P(c_f64), ctypes.c_int,
                                ctypes.c_int, c_f64, P(c_f64), P(PT2Stats)]
    L.fcg_plan_create_pair_screened.argtypes = [VP, VP, VP, VP, VP, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, P(c_f64), c_f64, P(VP)]
    L.fcg_plan_create_limited.argtypes = [VP, VP, VP, VP, ctypes.c_int, ctypes.c_int, c_i64,
                                          P(c_i64), P(VP)]
    L.fcg_plan_device_bytes.argtypes = [VP, P(c_i64)]
    L.fcg_set_plan_budget.argtypes = [c_i64]
    L.fcg_plan_apply_variant.argtypes = [VP, ctypes.c_int, ctypes.c_int, P(ctypes.c_int),
                                         P(ctypes.c_int)]
    L.fcg_plan_set_apply_variant.argtypes = [VP, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int]
    L.fcg_plan_budget.argtypes = [P(c_i64), P(c_i64)]
    L.fcg_apply_variant_counts.argtypes = [P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]
    L.fcg_set_build_class_cap.argtypes = [VP, ctypes.c_int]
    L.fcg_plan_create_matrix_free.argtypes = [VP, VP, VP, VP, ctypes.c_int, ctypes.c_int,
                                              P(VP)]
    L.fcg_serve_start.argtypes = [P(ServeOpts), P(VP)]
    L.fcg_serve_submit.argtypes = [VP, VP, VP, VP, ctypes.c_int, P(c_f64), P(c_i64)]
    L.fcg_serve_poll.argtypes = [VP, c_i64, ctypes.c_int, P(c_f64), P(c_f64), P(DavidsonStats),
                                 P(ServeTimes)]
    L.fcg_serve_stop.argtypes = [VP]
    L.fcg_tpsci_pt2_sparse.argtypes = [VP, VP, VP, VP, ctypes.c_int, P(c_f64), P(c_f64),
                                       ctypes.c_int, ctypes.c_int, c_f64, P(c_f64),
                                       P(PT2SparseStats)]
    L.fcg_tpsci_pt2_sparse_ex.argtypes = [VP, VP, VP, VP, ctypes.c_int, P(c_f64), P(c_f64),
                                          ctypes.c_int, ctypes.c_int, c_f64, ctypes.c_int,
                                          P(c_f64), P(PT2SparseStats)]
    L.fcg_open_sigma.argtypes = [VP, VP, VP, VP, VP, ctypes.c_int, ctypes.c_int, P(c_f64), c_f64,
                                 P(c_f64), P(ctypes.c_uint8)]
    L.fcg_open_sigma_sparse.argtypes = [VP, VP, VP, VP, ctypes.c_int, ctypes.c_int, P(c_f64),
                                        c_f64, P(VP)]
    L.fcg_open_result_size.argtypes = [VP, P(c_i64)]
    L.fcg_open_result_copy.argtypes = [VP, P(c_i16), P(c_i16), P(c_f64)]
    L.fcg_open_result_destroy.argtypes = [VP]
    L.fcg_open_result_destroy.restype = None
    L.fcg_plan_destroy.argtypes = [VP]
    L.fcg_plan_destroy.restype = None
    L.fcg_plan_stats.argtypes = [VP, P(c_f64)]
    L.fcg_sigma.argtypes = [VP, VP, ctypes.c_int, P(c_f64), P(c_f64), ctypes.c_int]
    L.fcg_sigma_device.argtypes = [VP, VP, ctypes.c_int, VP, VP, ctypes.c_int]
    L.fcg_sigma_device_noalloc.argtypes = [VP, VP, ctypes.c_int, VP, VP, ctypes.c_int]
    L.fcg_diagonal.argtypes = [VP, VP, P(c_f64)]
    L.fcg_space_diagonal.argtypes = [VP, VP, VP, VP, P(c_f64)]
    L.fcg_davidson.argtypes = [VP, VP, ctypes.c_int, P(c_f64), P(DavidsonOpts), P(c_f64),
                               P(c_f64), P(DavidsonStats)]
    L.fcg_last_sigma_time.argtypes = [VP, P(c_f64), P(ctypes.c_int)]
    L.fcg_event_record.argtypes = [VP, ctypes.c_int]
    L.fcg_event_elapsed.argtypes = [VP, ctypes.c_int, ctypes.c_int, P(c_f64)]
    L.fcg_synchronize.argtypes = [VP]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(P(t))


class Context:
    """One device context (stream, events, BLAS handle) of the library."""

    def __init__(self, device=0):
        self.L = load()
        h = VP()
        rc = self.L.fcg_create(device, ctypes.byref(h))
        if rc:
            raise FCGError("fcg_create(%d) failed: %s (is a GPU visible?)" % (device, ERRORS.get(rc, rc)))
        self.h = h

    def check(self, rc, what):
        if rc:
            msg = self.L.fcg_last_error(self.h).decode(errors="replace")
            raise FCGError("%s: %s %s" % (what, ERRORS.get(rc, rc), msg))

    def event(self, slot):
        self.check(self.L.fcg_event_record(self.h, slot), "fcg_event_record")

    def set_build_class_cap(self, cap):
        """Largest LDS class (3 tiny, 2 small, 1 middle, 0 full) of the block
        build for plans created on this context from now on; every class
        builds the same blocks bit for bit."""
        self.check(self.L.fcg_set_build_class_cap(self.h, int(cap)), "fcg_set_build_class_cap")

    def elapsed_ms(self, a, b):
        ms = c_f64()
        self.check(self.L.fcg_event_elapsed(self.h, a, b, ctypes.byref(ms)), "fcg_event_elapsed")
        return ms.value

    def synchronize(self):
        self.check(self.L.fcg_synchronize(self.h), "fcg_synchronize")

    def close(self):
        if getattr(self, "h", None):
            self.L.fcg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Handle:
    _destroy = None

    def __init__(self, ctx, h, keep=()):
        self.ctx, self.h, self._keep = ctx, h, keep

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                getattr(self.ctx.L, self._destroy)(self.h)
        except Exception:
            pass
        self.h = None


class DeviceOps(_Handle):
    _destroy = "fcg_ops_destroy"

    @classmethod
    def upload(cls, ctx, fops):
        gb = np.ascontiguousarray(fops.gblock, np.int32)
        off = np.ascontiguousarray(fops.gb_off, np.int64)
        dims = np.ascontiguousarray(fops.gb_dims, np.int32)
        data = np.ascontiguousarray(fops.gdata, np.float64)
        h = VP()
        ctx.check(ctx.L.fcg_ops_upload(ctx.h, fops.n_clusters, fops.S, fops.n_ops, _p(gb, c_i32),
                                       len(off), _p(off, c_i64), _p(dims, c_i32), _p(data, c_f64),
                                       data.size, ctypes.byref(h)), "fcg_ops_upload")
        return cls(ctx, h)


class DeviceHam(_Handle):
    _destroy = "fcg_ham_destroy"

    @classmethod
    def upload(cls, ctx, fham, n_clusters):
        types = {"nbody": np.int32, "clusters": np.int32, "ops": np.int32, "parity": np.int32,
                 "dims": np.int32, "delta": np.int16, "ioff": np.int64, "ints": np.float64,
                 "tc": np.int16, "tc_start": np.int32}
        a = {k: np.ascontiguousarray(getattr(fham, k), dtype=t) for k, t in types.items()}
        h = VP()
        ctx.check(ctx.L.fcg_ham_upload(
            ctx.h, n_clusters, fham.n_terms, _p(a["nbody"], c_i32), _p(a["clusters"], c_i32),
            _p(a["ops"], c_i32), _p(a["parity"], c_i32), _p(a["dims"], c_i32),
            _p(a["delta"], c_i16), _p(a["ioff"], c_i64), _p(a["ints"], c_f64), a["ints"].size,
            fham.n_tc, _p(a["tc"], c_i16), _p(a["tc_start"], c_i32), ctypes.byref(h)),
            "fcg_ham_upload")
        return cls(ctx, h)


class DeviceSpace(_Handle):
    _destroy = "fcg_space_destroy"

    @classmethod
    def tpsci(cls, ctx, state):
        fock = np.ascontiguousarray(state.fock, np.int16)
        fs = np.ascontiguousarray(state.fock_start, np.int64)
        cfg = np.ascontiguousarray(state.cfg, np.int16)
        h = VP()
        ctx.check(ctx.L.fcg_tpsci_space_upload(ctx.h, state.N, state.n_fock, _p(fock, c_i16),
                                               _p(fs, c_i64), len(state), _p(cfg, c_i16),
                                               ctypes.byref(h)), "fcg_tpsci_space_upload")
        return cls(ctx, h)

    @classmethod
    def bst(cls, ctx, flat):
        """`flat` is fermicg_amd.bst.FlatBST."""
        h = VP()
        ctx.check(ctx.L.fcg_bst_space_upload(
            ctx.h, flat.N, flat.n_fock, _p(flat.fock, c_i16), flat.n_box,
            _p(flat.box_fock, c_i32), _p(flat.lo, c_i16), _p(flat.hi, c_i16),
            _p(flat.rank, c_i16), _p(flat.foff, c_i64), _p(flat.factors, c_f64),
            flat.factors.size, ctypes.byref(h)), "fcg_bst_space_upload")
        return cls(ctx, h)

    def nvar(self):
        return int(self.ctx.L.fcg_space_nvar(self.h))


class Plan(_Handle):
    _destroy = "fcg_plan_destroy"

    @classmethod
    def create(cls, ctx, ham, ops, space, nbody=4, precisions=("fp64",)):
        mask = 0
        for p in precisions:
            mask |= 1 << PRECISION[p]
        h = VP()
        ctx.check(ctx.L.fcg_plan_create(ctx.h, ham.h, ops.h, space.h, nbody, mask,
                                        ctypes.byref(h)), "fcg_plan_create")
        pl = cls(ctx, h, keep=(ham, ops, space))
        pl.nvar = pl.nvar_ket = space.nvar()
        return pl

    @classmethod
    def create_matrix_free(cls, ctx, ham, ops, space, nbody=4, precisions=("fp64",)):
        """Matrix-free plan of a rank <= 1 TPSCI space (fcg_plan_create_matrix_free):
        sigma formed per call, nothing stored; FP64 and FP32."""
        mask = 0
        for p in precisions:
            mask |= 1 << PRECISION[p]
        h = VP()
        ctx.check(ctx.L.fcg_plan_create_matrix_free(ctx.h, ham.h, ops.h, space.h, nbody, mask,
                                                    ctypes.byref(h)),
                  "fcg_plan_create_matrix_free")
        pl = cls(ctx, h, keep=(ham, ops, space))
        pl.nvar = pl.nvar_ket = space.nvar()
        return pl

    @classmethod
    def create_limited(cls, ctx, ham, ops, space, max_bytes, nbody=4, precisions=("fp64",)):
        """Plan if its device footprint fits in max_bytes, else None (nothing
        allocated); returns (plan or None, footprint bytes)."""
        mask = 0
        for p in precisions:
            mask |= 1 << PRECISION[p]
        h = VP()
        need = c_i64()
        rc = ctx.L.fcg_plan_create_limited(ctx.h, ham.h, ops.h, space.h, nbody, mask,
                                           int(max_bytes), ctypes.byref(need), ctypes.byref(h))
        if rc == -2 and not h.value:
            return None, need.value
        ctx.check(rc, "fcg_plan_create_limited")
        pl = cls(ctx, h, keep=(ham, ops, space))
        pl.nvar = pl.nvar_ket = space.nvar()
        return pl, need.value

    def apply_variant(self, precision, R):
        """(multi-row, one-row) apply variants chosen for (precision, R); -1
        before the first sigma."""
        m, r = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.ctx.L.fcg_plan_apply_variant(self.h, PRECISION[precision], R,
                                                         ctypes.byref(m), ctypes.byref(r)),
                       "fcg_plan_apply_variant")
        return m.value, r.value

    def set_apply_variant(self, precision, R, multi, rows):
        """Pin the apply variants for (precision, R) instead of timing them."""
        self.ctx.check(self.ctx.L.fcg_plan_set_apply_variant(self.h, PRECISION[precision], R,
                                                             multi, rows),
                       "fcg_plan_set_apply_variant")

    def device_bytes(self):
        b = c_i64()
        self.ctx.check(self.ctx.L.fcg_plan_device_bytes(self.h, ctypes.byref(b)),
                       "fcg_plan_device_bytes")
        return b.value

    @classmethod
    def create_pair(cls, ctx, ham, ops, bra, ket, nbody=4, precisions=("fp64",), screen=None):
        """Plan of H between two spaces: sigma(bra) = H c(ket).  screen =
        (c, thresh) applies the open matvec's screening for ket state c."""
        mask = 0
        for p in precisions:
            mask |= 1 << PRECISION[p]
        h = VP()
        if screen is None:
            ctx.check(ctx.L.fcg_plan_create_pair(ctx.h, ham.h, ops.h, bra.h, ket.h, nbody, mask,
                                                 ctypes.byref(h)), "fcg_plan_create_pair")
        else:
            c, thresh = screen
            c = np.asarray(c, dtype=np.float64)
            if c.ndim == 1:
                c = c[:, None]
            cc = np.ascontiguousarray(c.T)
            ctx.check(ctx.L.fcg_plan_create_pair_screened(
                ctx.h, ham.h, ops.h, bra.h, ket.h, nbody, mask, c.shape[1], _p(cc, c_f64),
                float(thresh), ctypes.byref(h)), "fcg_plan_create_pair_screened")
        pl = cls(ctx, h, keep=(ham, ops, bra, ket))
        pl.nvar = bra.nvar()
        pl.nvar_ket = ket.nvar()
        return pl

    def stats(self):
        out = np.zeros(6)
        self.ctx.check(self.ctx.L.fcg_plan_stats(self.h, _p(out, c_f64)), "fcg_plan_stats")
        return {"pairs": int(out[0]), "tasks": int(out[1]), "block_elems": int(out[2]),
                "build_flops": out[3], "build_s": out[4] * 1e-6, "panel_cols": int(out[5])}

    def sigma(self, c, precision="fp64"):
        c = np.asarray(c, dtype=np.float64)
        one = c.ndim == 1
        if one:
            c = c[:, None]
        if c.shape[0] != self.nvar_ket:
            raise ValueError("vector length %d != space dimension %d" % (c.shape[0], self.nvar_ket))
        R = c.shape[1]
        cc = np.ascontiguousarray(c.T)  # (R, n) == column-major (n, R)
        out = np.zeros((R, self.nvar))
        self.ctx.check(self.ctx.L.fcg_sigma(self.ctx.h, self.h, R, _p(cc, c_f64), _p(out, c_f64),
                                            PRECISION[precision]), "fcg_sigma")
        s = out.T.copy()
        return s[:, 0] if one else s

    def sigma_device(self, c_ptr, s_ptr, R, precision="fp64"):
        self.ctx.check(self.ctx.L.fcg_sigma_device(self.ctx.h, self.h, R, VP(c_ptr), VP(s_ptr),
                                                   PRECISION[precision]), "fcg_sigma_device")

    def sigma_device_noalloc(self, c_ptr, s_ptr, R, precision="fp64", ctx=None):
        """fcg_sigma_device without any allocation (fixed device scratch;
        R x n_var <= 2^23); `ctx` defaults to the plan's context."""
        ctx = ctx or self.ctx
        ctx.check(ctx.L.fcg_sigma_device_noalloc(ctx.h, self.h, R, VP(c_ptr), VP(s_ptr),
                                                 PRECISION[precision]),
                  "fcg_sigma_device_noalloc")

    def last_sigma_ms(self):
        ms = c_f64()
        n = ctypes.c_int()
        self.ctx.check(self.ctx.L.fcg_last_sigma_time(self.ctx.h, ctypes.byref(ms), ctypes.byref(n)),
                       "fcg_last_sigma_time")
        return ms.value

    def diagonal(self):
        out = np.zeros(self.nvar)
        self.ctx.check(self.ctx.L.fcg_diagonal(self.ctx.h, self.h, _p(out, c_f64)), "fcg_diagonal")
        return out

    def davidson(self, v0, conv_thresh=1e-8, max_iter=200, max_ss_vecs=12, lindep_thresh=1e-12,
                 precision="fp64", refresh_k=0, precond=True, precond_diag=None):
        """Lowest R eigenpairs from start vectors v0 (n x R); precond_diag
        (n) replaces H's diagonal in the preconditioner."""
        v0 = np.asarray(v0, dtype=np.float64)
        if v0.ndim == 1:
            v0 = v0[:, None]
        R = v0.shape[1]
        vv = np.ascontiguousarray(v0.T)
        ev = np.zeros(R)
        vec = np.zeros((R, self.nvar))
        pd = None
        if precond_diag is not None:
            pd = np.ascontiguousarray(precond_diag, dtype=np.float64)
            if pd.shape != (self.nvar,):
                raise ValueError("precond_diag must have the space's length")
        o = DavidsonOpts(conv_thresh, max_iter, max_ss_vecs, lindep_thresh, PRECISION[precision],
                         refresh_k, 1 if precond else 0, _p(pd, c_f64) if pd is not None else None)
        st = DavidsonStats()
        rc = self.ctx.L.fcg_davidson(self.ctx.h, self.h, R, _p(vv, c_f64), ctypes.byref(o),
                                     _p(ev, c_f64), _p(vec, c_f64), ctypes.byref(st))
        if rc and rc != -6:
            self.ctx.check(rc, "fcg_davidson")
        stats = {"iterations": st.iterations, "sigma_fp64": st.sigma_fp64,
                 "sigma_low": st.sigma_low, "converged": bool(st.converged),
                 "residual": list(st.residual)[:R]}
        return ev, vec.T.copy(), stats


class Server:
    """The library's serving loop (fcg_serve_*): Davidson solves submitted as
    they arrive, run with continuous batching on `slots` contexts under an
    HBM budget for the live plans.  Handles passed to submit() stay
    referenced until their ticket is polled."""

    def __init__(self, slots=3, hbm_budget=2e8, precision="fp64", conv_thresh=1e-8, max_iter=300,
                 max_ss_vecs=12, lindep_thresh=1e-12, refresh_k=0, precond=True, nbody=4,
                 device=0):
        self.L = load()
        o = ServeOpts(device, slots, int(hbm_budget), nbody,
                      DavidsonOpts(conv_thresh, max_iter, max_ss_vecs, lindep_thresh,
                                   PRECISION[precision], refresh_k, 1 if precond else 0, None))
        h = VP()
        rc = self.L.fcg_serve_start(ctypes.byref(o), ctypes.byref(h))
        if rc:
            raise FCGError("fcg_serve_start: %s" % ERRORS.get(rc, rc))
        self.h = h
        self._live = {}

    def submit(self, ham, ops, space, v0):
        """Queue a solve from start vectors v0 (n x R); returns its ticket."""
        v0 = np.asarray(v0, dtype=np.float64)
        if v0.ndim == 1:
            v0 = v0[:, None]
        R = v0.shape[1]
        if v0.shape[0] != space.nvar():
            raise ValueError("start vectors do not match the space")
        vv = np.ascontiguousarray(v0.T)
        t = c_i64()
        rc = self.L.fcg_serve_submit(self.h, ham.h, ops.h, space.h, R, _p(vv, c_f64),
                                     ctypes.byref(t))
        if rc:
            raise FCGError("fcg_serve_submit: %s" % ERRORS.get(rc, rc))
        self._live[t.value] = (ham, ops, space, R, v0.shape[0])
        return t.value

    def poll(self, ticket, wait=True):
        """None while pending (wait=False); else a dict with the return code,
        energies, vectors (n x R), Davidson stats and times (seconds since
        the server started; plan_bytes)."""
        ham, ops, space, R, n = self._live[ticket]
        ev = np.zeros(R)
        vec = np.zeros((R, n))
        st = DavidsonStats()
        tm = ServeTimes()
        rc = self.L.fcg_serve_poll(self.h, ticket, 1 if wait else 0, _p(ev, c_f64),
                                   _p(vec, c_f64), ctypes.byref(st), ctypes.byref(tm))
        if rc == 1:
            return None
        del self._live[ticket]
        return {"rc": rc, "energies": ev, "vectors": vec.T.copy(),
                "stats": {"iterations": st.iterations, "sigma_fp64": st.sigma_fp64,
                          "sigma_low": st.sigma_low, "converged": bool(st.converged),
                          "residual": list(st.residual)[:R]},
                "times": {k: getattr(tm, k) for k, _ in ServeTimes._fields_}}

    def stop(self):
        if getattr(self, "h", None):
            self.L.fcg_serve_stop(self.h)
            self.h = None
            self._live.clear()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def tpsci_pt2(ctx, ham, ops, space, c, e0, nbody=4, h0_op=1, thresh=0.0):
    """Second-order energies e2[r] of the TPSCI state (space, c) over its
    first-order space (fcg_tpsci_pt2); returns (e2, stats dict)."""
    c = np.asarray(c, dtype=np.float64)
    if c.ndim == 1:
        c = c[:, None]
    R = c.shape[1]
    cc = np.ascontiguousarray(c.T)
    e0 = np.ascontiguousarray(np.broadcast_to(np.asarray(e0, np.float64), (R,)))
    e2 = np.zeros(R)
    st = PT2Stats()
    ctx.check(ctx.L.fcg_tpsci_pt2(ctx.h, ham.h, ops.h, space.h, R, _p(cc, c_f64), _p(e0, c_f64),
                                  nbody, h0_op, float(thresh), _p(e2, c_f64), ctypes.byref(st)),
              "fcg_tpsci_pt2")
    return e2, {k: getattr(st, k) for k, _ in PT2Stats._fields_}


def tpsci_pt2_sparse(ctx, ham, ops, space, c, e0, nbody=4, h0_op=1, thresh=0.0, precision="fp64"):
    """Second-order energies over the sparse first-order space
    (fcg_tpsci_pt2_sparse_ex; precision "fp64" or "fp32" for the
    contraction); returns (e2, stats dict)."""
    if precision not in ("fp64", "fp32"):
        raise ValueError("the sparse PT2 contracts in fp64 or fp32")
    c = np.asarray(c, dtype=np.float64)
    if c.ndim == 1:
        c = c[:, None]
    R = c.shape[1]
    cc = np.ascontiguousarray(c.T)
    e0 = np.ascontiguousarray(np.broadcast_to(np.asarray(e0, np.float64), (R,)))
    e2 = np.zeros(R)
    st = PT2SparseStats()
    ctx.check(ctx.L.fcg_tpsci_pt2_sparse_ex(ctx.h, ham.h, ops.h, space.h, R, _p(cc, c_f64),
                                            _p(e0, c_f64), nbody, h0_op, float(thresh),
                                            PRECISION[precision], _p(e2, c_f64),
                                            ctypes.byref(st)),
              "fcg_tpsci_pt2_sparse")
    return e2, {k: getattr(st, k) for k, _ in PT2SparseStats._fields_}


def open_sigma(ctx, ham, ops, bra, ket, c, nbody=4, thresh=0.0):
    """Screened open matvec onto the configurations of `bra` (fcg_open_sigma):
    returns (sigma [n_bra x R], hit [n_bra] bool, the configurations some
    element reached)."""
    c = np.asarray(c, dtype=np.float64)
    if c.ndim == 1:
        c = c[:, None]
    R = c.shape[1]
    if c.shape[0] != ket.nvar():
        raise ValueError("ket vector length %d != space dimension %d" % (c.shape[0], ket.nvar()))
    cc = np.ascontiguousarray(c.T)
    nb = bra.nvar()
    out = np.zeros((R, nb))
    hit = np.zeros(max(nb, 1), np.uint8)
    ctx.check(ctx.L.fcg_open_sigma(ctx.h, ham.h, ops.h, bra.h, ket.h, nbody, R, _p(cc, c_f64),
                                   float(thresh), _p(out, c_f64), _p(hit, ctypes.c_uint8)),
              "fcg_open_sigma")
    return out.T.copy(), hit[:nb].astype(bool)


def open_sigma_sparse(ctx, ham, ops, ket, c, n_clusters, nbody=4, thresh=0.0):
    """Screened open matvec generated per (ket configuration, term)
    (fcg_open_sigma_sparse): returns (fock [n, N, 2], cfg [n, N] 1-based,
    sigma [n, R]) of the configurations some element reached, in the
    engine's order."""
    c = np.asarray(c, dtype=np.float64)
    if c.ndim == 1:
        c = c[:, None]
    R = c.shape[1]
    if c.shape[0] != ket.nvar():
        raise ValueError("ket vector length %d != space dimension %d" % (c.shape[0], ket.nvar()))
    cc = np.ascontiguousarray(c.T)
    res = VP()
    ctx.check(ctx.L.fcg_open_sigma_sparse(ctx.h, ham.h, ops.h, ket.h, nbody, R, _p(cc, c_f64),
                                          float(thresh), ctypes.byref(res)),
              "fcg_open_sigma_sparse")
    try:
        n = c_i64(0)
        ctx.check(ctx.L.fcg_open_result_size(res, ctypes.byref(n)), "fcg_open_result_size")
        n = n.value
        fock = np.zeros((max(n, 1), n_clusters, 2), np.int16)
        cfg = np.zeros((max(n, 1), n_clusters), np.int16)
        sig = np.zeros((R, max(n, 1)))
        ctx.check(ctx.L.fcg_open_result_copy(res, _p(fock, c_i16), _p(cfg, c_i16),
                                             _p(sig, c_f64)), "fcg_open_result_copy")
    finally:
        ctx.L.fcg_open_result_destroy(res)
    return fock[:n], cfg[:n], sig[:, :n].T.copy()


def space_diagonal(ctx, ham, ops, space):
    """Diagonal of H over the configurations of a TPSCI space without a plan
    (fcg_space_diagonal)."""
    n = space.nvar()
    out = np.zeros(max(n, 1))
    ctx.check(ctx.L.fcg_space_diagonal(ctx.h, ham.h, ops.h, space.h, _p(out, c_f64)),
              "fcg_space_diagonal")
    return out[:n].copy()


def set_plan_budget(nbytes):
    """Process-wide device-byte budget for live plans (fcg_set_plan_budget)."""
    rc = load().fcg_set_plan_budget(int(nbytes))
    if rc:
        raise FCGError("fcg_set_plan_budget: %s" % ERRORS.get(rc, rc))


def plan_budget():
    """(budget, bytes held by live plans)."""
    b, h = c_i64(), c_i64()
    load().fcg_plan_budget(ctypes.byref(b), ctypes.byref(h))
    return b.value, h.value


def apply_variant_counts():
    """(multi-row variants, one-row variants, largest root count with all of them)."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = load().fcg_apply_variant_counts(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    if rc:
        raise FCGError("fcg_apply_variant_counts: %s" % ERRORS.get(rc, rc))
    return a.value, b.value, c.value

"""ctypes binding of the C ABI in ``include/bfe.h`` (``libbfe_hip.so``).

The library is the only compute path: every loader failure raises, there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libbfe_hip.so"

BFE_OK, BFE_ERR_ARG, BFE_ERR_HIP, BFE_ERR_NOMEM, BFE_ERR_NUMERIC, BFE_ERR_INTERNAL = range(6)
MODE_VEL_AT_CURRENT = 0
MODE_VEL_AT_LAST_HALF = 1
PREC_DOUBLE = 0
PREC_SINGLE = 1
TUNE_MULTI_ROWS_PER_BLOCK = 1
TUNE_XCD_REMAP = 2
TUNE_OPERATOR = 3
OPERATOR_CSR = 0
OPERATOR_EBE = 1
OPERATOR_NODE = 2
TUNE_EBE_GROUP = 4
TUNE_EBE_KERNEL = 5
TUNE_MULTI_ROWS_PER_WAVE = 6
TUNE_GRAPH = 7
TUNE_MULTI_GROUP = 8
TUNE_PERSISTENT = 9
TUNE_PERSISTENT_GRID = 10
TUNE_MULTI_OPERATOR = 11
TUNE_MULTI_NODES_PER_WAVE = 12
TUNE_MULTI_SLIDE = 13
TUNE_ADJOINT_PAIRS = 14
# per-element field types of bfe_run_field
FIELD_KE, FIELD_MEAN, FIELD_DIV, FIELD_CURL = 0, 1, 2, 3

# Every symbol the header declares.
EXPORTS = (
    "bfe_create", "bfe_run", "bfe_run_batched", "bfe_set_gradient_elements",
    "bfe_run_checkpointed", "bfe_run_checkpointed_batched", "bfe_adjoint", "bfe_set_tuning", "bfe_set_element_operator",
    "bfe_set_element_groups", "bfe_set_field_elements", "bfe_run_field_ke",
    "bfe_set_field_geometry", "bfe_run_field",
    "bfe_set_kernel_timing",
    "bfe_info_get",
    "bfe_last_run_stats", "bfe_destroy", "bfe_last_error", "bfe_assemble_pattern",
    "bfe_assemble_fill", "bfe_assemble_last_error", "bfe_assemble_device",
    "bfe_assemble_device_fill", "bfe_assemble_device_free", "bfe_convolve",
    "bfe_disps_to_forces", "bfe_disps_to_forces_noalloc", "bfe_element_templates", "bfe_element_centres",
    "bfe_points_in_polygon", "bfe_dist_points_to_polygon", "bfe_tri_grid",
    "bfe_remove_unused_nodes", "bfe_get_edges",
)


class BfeInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("nnz_k", C.c_int64), ("nnz_sell", C.c_int64),
        ("n_interface", C.c_int64), ("nnz_interface", C.c_int64), ("nnz_xinv", C.c_int64),
        ("damped", C.c_int32), ("mode", C.c_int32), ("precision", C.c_int32),
        ("device", C.c_int32), ("device_bytes", C.c_int64), ("operator_kind", C.c_int32),
        ("ebe_slots", C.c_int32), ("ebe_templates", C.c_int32), ("ebe_pattern", C.c_int32),
        ("ebe_elements", C.c_int64), ("ebe_check", C.c_double), ("ebe_classes", C.c_int32),
        ("ebe_groups", C.c_int32), ("multi_group", C.c_int32), ("ebe_sparse", C.c_int32),
        ("multi_operator", C.c_int32), ("grad_pairs", C.c_int32),
    ]


class BfeElementGroup(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64), ("dpn", C.c_int32), ("npe", C.c_int32), ("n_el", C.c_int64),
        ("el_nodes", C.c_void_p), ("el_tmpl", C.c_void_p), ("el_scale", C.c_void_p),
        ("n_tmpl", C.c_int32), ("reserved", C.c_int32), ("tmpl", C.c_void_p),
        ("node_dof", C.c_void_p),
    ]


class BfeRunStats(C.Structure):
    _fields_ = [
        ("loop_ms", C.c_double), ("steps", C.c_int64), ("launches", C.c_int64),
        ("columns", C.c_int64), ("nonfinite", C.c_int64), ("step_kernel_ms", C.c_double),
        ("step_kernel_samples", C.c_int64),
    ]


class BfeError(RuntimeError):
    """Raised for any non-zero bfe_status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"bfe status {status}: {message}")
        self.status = status


_lib = None

_p = C.POINTER
_i64p = _p(C.c_int64)
_i32p = _p(C.c_int32)
_f64p = _p(C.c_double)


def load() -> C.CDLL:
    """Loads ``libbfe_hip.so``; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python bristolfe-v2_amd/build.py or __graft_entry__.build())")
    lib = C.CDLL(str(LIB_PATH))
    lib.bfe_create.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _i64p, _i32p, _f64p,
                               C.c_double, C.c_int32, C.c_int32, C.c_int32, _p(C.c_void_p)]
    lib.bfe_create.restype = C.c_int
    lib.bfe_run.argtypes = [C.c_void_p, C.c_int32, _i32p, C.c_int32, _f64p, C.c_int32, _i32p,
                            C.c_int32, _f64p, C.c_int32, C.c_int32, _f64p, _i32p]
    lib.bfe_run.restype = C.c_int
    lib.bfe_run_batched.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _i32p, _i32p, C.c_int32,
                                    _f64p, C.c_int32, _i32p, C.c_int32, _f64p, _i32p]
    lib.bfe_run_batched.restype = C.c_int
    lib.bfe_set_gradient_elements.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _i32p, _f64p,
                                              _f64p, _f64p]
    lib.bfe_set_gradient_elements.restype = C.c_int
    lib.bfe_run_checkpointed.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _i32p, C.c_int32, _f64p,
                                         C.c_int32, _i32p, C.c_int32, _f64p, C.c_int32, _i32p]
    lib.bfe_run_checkpointed.restype = C.c_int
    lib.bfe_run_checkpointed_batched.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                                 _i32p, _i32p, C.c_int32, _f64p, C.c_int32, _i32p,
                                                 C.c_int32, _f64p, C.c_int32, _i32p]
    lib.bfe_run_checkpointed_batched.restype = C.c_int
    lib.bfe_adjoint.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p]
    lib.bfe_adjoint.restype = C.c_int
    lib.bfe_set_tuning.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.bfe_set_tuning.restype = C.c_int
    lib.bfe_set_element_operator.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int64,
                                             C.c_int32, _i32p, _i32p, _f64p, C.c_int32, _f64p,
                                             C.c_double]
    lib.bfe_set_element_operator.restype = C.c_int
    lib.bfe_set_element_groups.argtypes = [C.c_void_p, C.c_int32, _p(BfeElementGroup), C.c_double]
    lib.bfe_set_element_groups.restype = C.c_int
    lib.bfe_set_field_elements.argtypes = [C.c_void_p, C.c_int64, C.c_int32, _i32p, C.c_int64,
                                           C.c_int32, _i32p]
    lib.bfe_set_field_elements.restype = C.c_int
    lib.bfe_run_field_ke.argtypes = [C.c_void_p, C.c_int32, _i32p, C.c_int32, _f64p, C.c_int32,
                                     _i32p, C.c_int32, _f64p, C.c_int32, _f64p, _i32p]
    lib.bfe_run_field_ke.restype = C.c_int
    lib.bfe_set_field_geometry.argtypes = [C.c_void_p, C.c_int32, _f64p]
    lib.bfe_set_field_geometry.restype = C.c_int
    lib.bfe_run_field.argtypes = [C.c_void_p, C.c_int32, _i32p, C.c_int32, _f64p, C.c_int32,
                                  _i32p, C.c_int32, _f64p, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32, _f64p, _i32p]
    lib.bfe_run_field.restype = C.c_int
    lib.bfe_set_kernel_timing.argtypes = [C.c_void_p, C.c_int32]
    lib.bfe_set_kernel_timing.restype = C.c_int
    lib.bfe_info_get.argtypes = [C.c_void_p, _p(BfeInfo)]
    lib.bfe_info_get.restype = C.c_int
    lib.bfe_last_run_stats.argtypes = [C.c_void_p, _p(BfeRunStats)]
    lib.bfe_last_run_stats.restype = C.c_int
    lib.bfe_destroy.argtypes = [C.c_void_p]
    lib.bfe_destroy.restype = C.c_int
    lib.bfe_last_error.argtypes = []
    lib.bfe_last_error.restype = C.c_char_p
    lib.bfe_assemble_pattern.argtypes = [C.c_int64, C.c_int32, _f64p, C.c_int64, C.c_int32, _i32p,
                                         _i32p, _i32p, _f64p, C.c_int32, _f64p, _f64p, C.c_int32,
                                         C.c_double, C.c_double, C.c_double, _i64p, _i64p, _i32p,
                                         _p(C.c_void_p)]
    lib.bfe_assemble_pattern.restype = C.c_int
    lib.bfe_assemble_fill.argtypes = [C.c_void_p, _i64p, _i32p, _f64p, _f64p, _f64p, _i64p]
    lib.bfe_assemble_fill.restype = C.c_int
    lib.bfe_assemble_last_error.argtypes = []
    lib.bfe_assemble_last_error.restype = C.c_char_p
    lib.bfe_assemble_device.argtypes = [C.c_int32, C.c_int64, C.c_int32, _f64p, C.c_int64, C.c_int32,
                                        _i32p, _i32p, _i32p, _f64p, C.c_int32, _f64p, _f64p,
                                        C.c_int32, C.c_double, C.c_double, C.c_double, _i64p, _i64p,
                                        _i64p, _i32p, _p(C.c_void_p)]
    lib.bfe_assemble_device.restype = C.c_int
    lib.bfe_assemble_device_fill.argtypes = [C.c_void_p, _i64p, _i32p, _f64p, _i64p, _i32p, _f64p,
                                             _f64p, _i64p]
    lib.bfe_assemble_device_fill.restype = C.c_int
    lib.bfe_assemble_device_free.argtypes = [C.c_void_p]
    lib.bfe_assemble_device_free.restype = C.c_int
    lib.bfe_convolve.argtypes = [C.c_int32, C.c_int64, C.c_int32, _f64p, C.c_int32, _f64p, C.c_int32,
                                 _f64p]
    lib.bfe_convolve.restype = C.c_int
    lib.bfe_disps_to_forces.argtypes = [C.c_int32, C.c_int64, _i64p, _i32p, _f64p, _i64p, _i32p,
                                        _f64p, _f64p, C.c_double, C.c_int32, _f64p, _i32p,
                                        C.c_int32, C.c_int32, C.c_int64, _i32p, _f64p]
    lib.bfe_disps_to_forces.restype = C.c_int
    lib.bfe_disps_to_forces_noalloc.argtypes = lib.bfe_disps_to_forces.argtypes
    lib.bfe_disps_to_forces_noalloc.restype = C.c_int
    lib.bfe_element_templates.argtypes = [C.c_int32, C.c_int64, C.c_int32, _f64p, C.c_int64,
                                          C.c_int32, _i32p, _i32p, _i64p, C.c_int32, _i32p, _i64p,
                                          _i32p, _f64p]
    lib.bfe_element_templates.restype = C.c_int
    lib.bfe_element_centres.argtypes = [C.c_int64, C.c_int32, _f64p, C.c_int64, C.c_int32, _i64p,
                                        _f64p]
    lib.bfe_element_centres.restype = C.c_int
    lib.bfe_points_in_polygon.argtypes = [C.c_int64, _f64p, C.c_int32, _f64p, _f64p,
                                          _p(C.c_uint8)]
    lib.bfe_points_in_polygon.restype = C.c_int
    lib.bfe_dist_points_to_polygon.argtypes = [C.c_int64, _f64p, C.c_int32, _f64p, _f64p]
    lib.bfe_dist_points_to_polygon.restype = C.c_int
    lib.bfe_tri_grid.argtypes = [C.c_int32, C.c_int64, C.c_int64, _f64p, _f64p, C.c_double, _f64p,
                                 _i64p, _i64p]
    lib.bfe_tri_grid.restype = C.c_int
    lib.bfe_remove_unused_nodes.argtypes = [C.c_int64, C.c_int64, C.c_int32, _i64p, _i64p, _i64p]
    lib.bfe_remove_unused_nodes.restype = C.c_int
    lib.bfe_get_edges.argtypes = [C.c_int64, C.c_int32, _i64p, _i64p, _i64p, _i64p, _i64p]
    lib.bfe_get_edges.restype = C.c_int
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != BFE_OK:
        msg = load().bfe_last_error()
        raise BfeError(status, msg.decode() if msg else "")


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(_p(ctype))


class Plan:
    """Owns one ``bfe_handle``: operator, coefficients and a HIP stream."""

    def __init__(self, n, k_ptr, k_col, k_val, m, c_ptr, c_col, c_val, dt, mode, precision,
                 device=0):
        lib = load()
        self._lib = lib
        self.n = int(n)
        arrs = dict(
            k_ptr=np.ascontiguousarray(k_ptr, np.int64), k_col=np.ascontiguousarray(k_col, np.int32),
            k_val=np.ascontiguousarray(k_val, np.float64), m=np.ascontiguousarray(m, np.float64),
        )
        if c_ptr is not None:
            arrs.update(c_ptr=np.ascontiguousarray(c_ptr, np.int64),
                        c_col=np.ascontiguousarray(c_col, np.int32),
                        c_val=np.ascontiguousarray(c_val, np.float64))
        else:
            arrs.update(c_ptr=None, c_col=None, c_val=None)
        if arrs["m"].shape[0] != self.n or arrs["k_ptr"].shape[0] != self.n + 1:
            raise ValueError("inconsistent sizes for K / m")
        h = C.c_void_p()
        check(lib.bfe_create(
            self.n, ptr(arrs["k_ptr"], C.c_int64), ptr(arrs["k_col"], C.c_int32),
            ptr(arrs["k_val"], C.c_double), ptr(arrs["m"], C.c_double),
            ptr(arrs["c_ptr"], C.c_int64), ptr(arrs["c_col"], C.c_int32),
            ptr(arrs["c_val"], C.c_double), float(dt), int(mode), int(precision), int(device),
            C.byref(h)))
        self._h = h

    @classmethod
    def from_scipy(cls, K, m, C_mat, dt, mode, precision, device=0):
        K = K.tocsr()
        c_args = (None, None, None)
        if C_mat is not None and C_mat.nnz > 0:
            Cc = C_mat.tocsr()
            c_args = (Cc.indptr, Cc.indices, Cc.data)
        return cls(K.shape[0], K.indptr, K.indices, K.data, m, *c_args, dt, mode, precision, device)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(self._lib.bfe_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        inf = BfeInfo()
        check(self._lib.bfe_info_get(self._h, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in BfeInfo._fields_}

    def set_tuning(self, key: int, value: int) -> None:
        check(self._lib.bfe_set_tuning(self._h, int(key), int(value)))

    def set_kernel_timing(self, every: int) -> None:
        check(self._lib.bfe_set_kernel_timing(self._h, int(every)))

    def stats(self) -> dict:
        st = BfeRunStats()
        check(self._lib.bfe_last_run_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in BfeRunStats._fields_}

    def run(self, n_t, frc_idx, frcs, hist_idx, field_every=0, field_is_vel=1):
        """0-based indices; frcs is (rows x n_t). Returns (hist, field, ti_start)."""
        frc_idx = np.ascontiguousarray(frc_idx, np.int32).ravel()
        hist_idx = np.ascontiguousarray(hist_idx, np.int32).ravel()
        p, q = frc_idx.size, hist_idx.size
        frcs = np.asarray(frcs, np.float64)
        if frcs.ndim == 1:
            frcs = frcs[None, :]
        rows = frcs.shape[0] if p > 0 else 0
        Fcol = np.asfortranarray(frcs) if p > 0 else np.zeros((1, 1))
        hist = np.empty((q, n_t), order="F")  # the library writes every entry
        n_frames = (n_t - 1) // field_every + 1 if field_every and field_every > 0 else 0
        fld = np.empty((self.n, n_frames), order="F") if n_frames else None
        ti = C.c_int32()
        check(self._lib.bfe_run(
            self._h, int(n_t), ptr(frc_idx, C.c_int32), p, Fcol.ctypes.data_as(_f64p), rows,
            ptr(hist_idx, C.c_int32), q, hist.ctypes.data_as(_f64p),
            int(field_every if n_frames else 0), int(field_is_vel),
            fld.ctypes.data_as(_f64p) if fld is not None else None, C.byref(ti)))
        return hist, fld, int(ti.value)

    def set_element_operator(self, op, check_tol=0.0):
        """Attach an element-by-element operator (``ElementOperator`` or
        ``ElementGroups`` of bristolfe_v2_amd.ebe)."""
        if len(op.groups) > 1 or op.groups[0].node_dof is not None:
            self.set_element_groups(op.groups, check_tol)
            return
        el_nodes = np.ascontiguousarray(op.el_nodes, np.int32)
        n_el, npe = el_nodes.shape
        el_tmpl = np.ascontiguousarray(op.el_tmpl, np.int32).ravel()
        el_scale = np.ascontiguousarray(op.el_scale, np.float64).ravel()
        tmpl = np.ascontiguousarray(op.tmpl, np.float64)
        check(self._lib.bfe_set_element_operator(
            self._h, int(op.n_nodes), int(op.dpn), n_el, npe, ptr(el_nodes, C.c_int32),
            ptr(el_tmpl, C.c_int32), ptr(el_scale, C.c_double), int(tmpl.shape[0]),
            ptr(tmpl, C.c_double), float(check_tol)))

    def set_field_elements(self, els, gl_lookup):
        """Mesh tables of the device 'KE' field: els (n_el, npe) and gl_lookup
        (n_nodes, max_df) 1-based with 0 = absent, as the model holds them."""
        els0 = np.ascontiguousarray(np.asarray(els, np.int64) - 1, np.int32)
        nd0 = np.ascontiguousarray(np.asarray(gl_lookup, np.int64) - 1, np.int32)
        self.n_field_el = els0.shape[0]
        self.field_dim = None  # new tables drop the node coordinates
        check(self._lib.bfe_set_field_elements(self._h, nd0.shape[0], nd0.shape[1],
                                               ptr(nd0, C.c_int32), els0.shape[0], els0.shape[1],
                                               ptr(els0, C.c_int32)))

    def set_field_geometry(self, nds):
        """Node coordinates (n_nodes, dim) for the 'div' and 'curl' fields."""
        nds = np.ascontiguousarray(nds, np.float64)
        check(self._lib.bfe_set_field_geometry(self._h, int(nds.shape[1]), ptr(nds, C.c_double)))
        self.field_dim = int(nds.shape[1])

    def run_field(self, n_t, frc_idx, frcs, hist_idx, field_every, field_type, field_is_vel,
                  comp=0):
        """bfe_run with a per-element field reduced on the device (FIELD_KE,
        FIELD_MEAN of DOF column comp, FIELD_DIV, FIELD_CURL, from velocity
        or displacement frames): returns (hist, el (n_el x n_frames),
        ti_start)."""
        return self._run_field(n_t, frc_idx, frcs, hist_idx, field_every,
                               (int(field_type), 1 if field_is_vel else 0, int(comp)))

    def run_field_ke(self, n_t, frc_idx, frcs, hist_idx, field_every):
        """bfe_run with the 'KE' field reduced on the device: returns (hist,
        el_ke (n_el x n_frames), ti_start)."""
        return self._run_field(n_t, frc_idx, frcs, hist_idx, field_every, None)

    def _run_field(self, n_t, frc_idx, frcs, hist_idx, field_every, kind):
        frc_idx = np.ascontiguousarray(frc_idx, np.int32).ravel()
        hist_idx = np.ascontiguousarray(hist_idx, np.int32).ravel()
        p, q = frc_idx.size, hist_idx.size
        frcs = np.asarray(frcs, np.float64)
        if frcs.ndim == 1:
            frcs = frcs[None, :]
        rows = frcs.shape[0] if p > 0 else 0
        Fcol = np.asfortranarray(frcs) if p > 0 else np.zeros((1, 1))
        hist = np.empty((q, n_t), order="F")  # the library writes every entry
        n_frames = (n_t - 1) // int(field_every) + 1
        el = np.empty((self.n_field_el, n_frames), order="F")
        ti = C.c_int32()
        if kind is None:
            check(self._lib.bfe_run_field_ke(
                self._h, int(n_t), ptr(frc_idx, C.c_int32), p, Fcol.ctypes.data_as(_f64p), rows,
                ptr(hist_idx, C.c_int32), q, hist.ctypes.data_as(_f64p), int(field_every),
                el.ctypes.data_as(_f64p), C.byref(ti)))
        else:
            check(self._lib.bfe_run_field(
                self._h, int(n_t), ptr(frc_idx, C.c_int32), p, Fcol.ctypes.data_as(_f64p), rows,
                ptr(hist_idx, C.c_int32), q, hist.ctypes.data_as(_f64p), int(field_every),
                kind[0], kind[1], kind[2], el.ctypes.data_as(_f64p), C.byref(ti)))
        return hist, el, int(ti.value)

    def set_element_groups(self, groups, check_tol=0.0):
        keep = []
        arr = (BfeElementGroup * len(groups))()
        for i, g in enumerate(groups):
            el_nodes = np.ascontiguousarray(g.el_nodes, np.int32)
            el_tmpl = np.ascontiguousarray(g.el_tmpl, np.int32).ravel()
            el_scale = np.ascontiguousarray(g.el_scale, np.float64).ravel()
            tmpl = np.ascontiguousarray(g.tmpl, np.float64)
            nd = None if g.node_dof is None else np.ascontiguousarray(g.node_dof, np.int32)
            keep += [el_nodes, el_tmpl, el_scale, tmpl, nd]
            arr[i] = BfeElementGroup(
                int(g.n_nodes), int(g.dpn), int(el_nodes.shape[1]), int(el_nodes.shape[0]),
                el_nodes.ctypes.data, el_tmpl.ctypes.data, el_scale.ctypes.data,
                int(tmpl.shape[0]), 0, tmpl.ctypes.data, None if nd is None else nd.ctypes.data)
        check(self._lib.bfe_set_element_groups(self._h, len(groups), arr, float(check_tol)))
        del keep

    def set_gradient_elements(self, el_dofs, el_K, el_m, el_mu):
        """el_dofs (n_el, k) 0-based (-1 unused), el_K (n_el, k, k), el_m (n_el, k), el_mu (n_el)."""
        el_dofs = np.ascontiguousarray(el_dofs, np.int32)
        n_el, k = el_dofs.shape
        el_K = np.ascontiguousarray(el_K, np.float64).reshape(n_el, k, k)
        el_m = np.ascontiguousarray(el_m, np.float64).reshape(n_el, k)
        el_mu = np.ascontiguousarray(el_mu, np.float64).reshape(n_el)
        self.n_el = n_el
        check(self._lib.bfe_set_gradient_elements(self._h, n_el, k, ptr(el_dofs, C.c_int32),
                                                  ptr(el_K, C.c_double), ptr(el_m, C.c_double),
                                                  ptr(el_mu, C.c_double)))

    def run_checkpointed(self, slot, n_t, frc_idx, frcs, hist_idx, ckpt_every):
        frc_idx = np.ascontiguousarray(frc_idx, np.int32).ravel()
        hist_idx = np.ascontiguousarray(hist_idx, np.int32).ravel()
        p, q = frc_idx.size, hist_idx.size
        frcs = np.asarray(frcs, np.float64)
        if frcs.ndim == 1:
            frcs = frcs[None, :]
        rows = frcs.shape[0] if p > 0 else 0
        Fcol = np.asfortranarray(frcs) if p > 0 else np.zeros((1, 1))
        hist = np.empty((q, n_t), order="F")  # the library writes every entry
        ti = C.c_int32()
        check(self._lib.bfe_run_checkpointed(
            self._h, int(slot), int(n_t), ptr(frc_idx, C.c_int32), p, Fcol.ctypes.data_as(_f64p),
            rows, ptr(hist_idx, C.c_int32), q, hist.ctypes.data_as(_f64p), int(ckpt_every),
            C.byref(ti)))
        return hist, int(ti.value)

    def run_checkpointed_batched(self, slot, n_t, n_src, frc_idx, frc_src, frcs, hist_idx,
                                 ckpt_every):
        """bfe_run_checkpointed_batched; returns hist (q, n_src, n_t) and ti_start."""
        frc_idx = np.ascontiguousarray(frc_idx, np.int32).ravel()
        frc_src = np.ascontiguousarray(frc_src, np.int32).ravel()
        hist_idx = np.ascontiguousarray(hist_idx, np.int32).ravel()
        p, q = frc_idx.size, hist_idx.size
        frcs = np.asarray(frcs, np.float64)
        if frcs.ndim == 1:
            frcs = frcs[None, :]
        rows = frcs.shape[0] if p > 0 else 0
        Fcol = np.asfortranarray(frcs) if p > 0 else np.zeros((1, 1))
        hist = np.empty((q, n_src, n_t), order="F")  # the library writes every entry
        ti = C.c_int32()
        check(self._lib.bfe_run_checkpointed_batched(
            self._h, int(slot), int(n_t), int(n_src), ptr(frc_idx, C.c_int32),
            ptr(frc_src, C.c_int32), p, Fcol.ctypes.data_as(_f64p), rows,
            ptr(hist_idx, C.c_int32), q, hist.ctypes.data_as(_f64p), int(ckpt_every),
            C.byref(ti)))
        return hist, int(ti.value)

    def adjoint(self, slot, hist_grad):
        g = np.asfortranarray(hist_grad, dtype=np.float64)
        gk = np.zeros(self.n_el)
        gm = np.zeros(self.n_el)
        check(self._lib.bfe_adjoint(self._h, int(slot), g.ctypes.data_as(_f64p),
                                    gk.ctypes.data_as(_f64p), gm.ctypes.data_as(_f64p)))
        return gk, gm

    def run_batched(self, n_t, n_src, frc_idx, frc_src, frcs, hist_idx):
        """Returns hist of shape (q, n_src, n_t) and ti_start."""
        frc_idx = np.ascontiguousarray(frc_idx, np.int32).ravel()
        frc_src = np.ascontiguousarray(frc_src, np.int32).ravel()
        hist_idx = np.ascontiguousarray(hist_idx, np.int32).ravel()
        p, q = frc_idx.size, hist_idx.size
        frcs = np.asarray(frcs, np.float64)
        if frcs.ndim == 1:
            frcs = frcs[None, :]
        rows = frcs.shape[0] if p > 0 else 0
        Fcol = np.asfortranarray(frcs) if p > 0 else np.zeros((1, 1))
        hist = np.empty((q, n_src, n_t), order="F")  # the library writes every entry
        ti = C.c_int32()
        check(self._lib.bfe_run_batched(
            self._h, int(n_t), int(n_src), ptr(frc_idx, C.c_int32), ptr(frc_src, C.c_int32), p,
            Fcol.ctypes.data_as(_f64p), rows, ptr(hist_idx, C.c_int32), q,
            hist.ctypes.data_as(_f64p), C.byref(ti)))
        return hist, int(ti.value)


def assemble_native(nds, els0, el_type, el_mat0, el_abs, mat_D, mat_rho, dof_mask, p_law,
                    max_damping, msr):
    """Calls bfe_assemble_pattern/fill. 0-based element nodes and materials.
    Returns (rowptr, col, kval, m, c, gl_lookup)."""
    lib = load()
    nds = np.ascontiguousarray(nds, np.float64)
    els0 = np.ascontiguousarray(els0, np.int32)
    el_type = np.ascontiguousarray(el_type, np.int32)
    el_mat0 = np.ascontiguousarray(el_mat0, np.int32)
    el_abs = np.ascontiguousarray(el_abs, np.float64)
    mat_D = np.ascontiguousarray(mat_D, np.float64)
    mat_rho = np.ascontiguousarray(mat_rho, np.float64)
    n_dof, nnz, max_df, work = C.c_int64(), C.c_int64(), C.c_int32(), C.c_void_p()
    st = lib.bfe_assemble_pattern(
        nds.shape[0], nds.shape[1], ptr(nds, C.c_double), els0.shape[0], els0.shape[1],
        ptr(els0, C.c_int32), ptr(el_type, C.c_int32), ptr(el_mat0, C.c_int32), ptr(el_abs, C.c_double),
        mat_rho.size, ptr(mat_D, C.c_double), ptr(mat_rho, C.c_double), int(dof_mask), float(p_law),
        float(max_damping), float(msr), C.byref(n_dof), C.byref(nnz), C.byref(max_df), C.byref(work))
    if st != BFE_OK:
        raise BfeError(st, lib.bfe_assemble_last_error().decode())
    n, z = n_dof.value, nnz.value
    rowptr = np.empty(n + 1, np.int64)
    col = np.empty(z, np.int32)
    kval = np.empty(z, np.float64)
    m = np.empty(n, np.float64)
    c = np.empty(n, np.float64)
    gl = np.empty((nds.shape[0], max_df.value), np.int64)
    st = lib.bfe_assemble_fill(work, ptr(rowptr, C.c_int64), ptr(col, C.c_int32), ptr(kval, C.c_double),
                               ptr(m, C.c_double), ptr(c, C.c_double), ptr(gl, C.c_int64))
    if st != BFE_OK:
        raise BfeError(st, lib.bfe_assemble_last_error().decode())
    return rowptr, col, kval, m, c, gl


def element_centres(nds, els):
    """bfe_element_centres: els 1-based (0 = absent), as the model holds it."""
    lib = load()
    nds = np.ascontiguousarray(nds, np.float64)
    els = np.ascontiguousarray(els, np.int64)
    if els.ndim == 1:
        els = els[:, None]
    out = np.empty((els.shape[0], nds.shape[1]), np.float64)
    check(lib.bfe_element_centres(nds.shape[0], nds.shape[1], ptr(nds, C.c_double), els.shape[0],
                                  els.shape[1], ptr(els, C.c_int64), ptr(out, C.c_double)))
    return out


def points_in_polygon(pts, vx, vy):
    lib = load()
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 2)
    vx = np.ascontiguousarray(vx, np.float64).ravel()
    vy = np.ascontiguousarray(vy, np.float64).ravel()
    out = np.empty(pts.shape[0], np.uint8)
    check(lib.bfe_points_in_polygon(pts.shape[0], ptr(pts, C.c_double), vx.size,
                                    ptr(vx, C.c_double), ptr(vy, C.c_double),
                                    ptr(out, C.c_uint8)))
    return out.astype(bool)


def dist_points_to_polygon(pts, bdry):
    lib = load()
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 2)
    bdry = np.ascontiguousarray(bdry, np.float64).reshape(-1, 2)
    out = np.empty(pts.shape[0], np.float64)
    check(lib.bfe_dist_points_to_polygon(pts.shape[0], ptr(pts, C.c_double), bdry.shape[0],
                                         ptr(bdry, C.c_double), ptr(out, C.c_double)))
    return out


def tri_grid(iso, x, y, shift):
    """bfe_tri_grid: (nds, els) of the structured triangle grid on axes x, y
    (els 1-based, the reference's block order), before any region clip."""
    lib = load()
    x = np.ascontiguousarray(x, np.float64).ravel()
    y = np.ascontiguousarray(y, np.float64).ravel()
    n_el = C.c_int64(0)
    check(lib.bfe_tri_grid(int(iso), x.size, y.size, ptr(x, C.c_double), ptr(y, C.c_double),
                           float(shift), None, None, C.byref(n_el)))
    nds = np.empty((x.size * y.size, 2), np.float64)
    els = np.empty((n_el.value, 3), np.int64)
    check(lib.bfe_tri_grid(int(iso), x.size, y.size, ptr(x, C.c_double), ptr(y, C.c_double),
                           float(shift), ptr(nds, C.c_double), ptr(els, C.c_int64),
                           C.byref(n_el)))
    return nds, els


def remove_unused_nodes(n_nodes, els):
    """bfe_remove_unused_nodes: (renumbered els, old_nds) with old_nds 1-based."""
    lib = load()
    els = np.array(els, dtype=np.int64, order="C", copy=True)
    if els.ndim == 1:
        els = els[:, None]
    old = np.empty(max(int(n_nodes), 1), np.int64)
    nu = C.c_int64(0)
    check(lib.bfe_remove_unused_nodes(int(n_nodes), els.shape[0], els.shape[1],
                                      ptr(els, C.c_int64), ptr(old, C.c_int64), C.byref(nu)))
    return els, old[:nu.value].copy()


def get_edges(els):
    """bfe_get_edges: (ed, ptr, el) with ed (n_ed x 2) the unique sorted edges
    and el[ptr[i]:ptr[i+1]] the elements (1-based, largest first) on edge i."""
    lib = load()
    els = np.ascontiguousarray(els, np.int64)
    if els.ndim == 1:
        els = els[:, None]
    m = max(els.shape[0] * els.shape[1], 1)
    ed = np.empty((m, 2), np.int64)
    pt = np.empty(m + 1, np.int64)
    el = np.empty(m, np.int64)
    ne = C.c_int64(0)
    check(lib.bfe_get_edges(els.shape[0], els.shape[1], ptr(els, C.c_int64), C.byref(ne),
                            ptr(ed, C.c_int64), ptr(pt, C.c_int64), ptr(el, C.c_int64)))
    n = ne.value
    return ed[:n].copy(), pt[:n + 1].copy(), el[:pt[n]].copy()


def element_templates(nds, els0, mat, mult, max_templates, with_jfactor, device=0):
    """bfe_element_templates: returns (el_tmpl, first, jfactor or None), or
    None when there are more than max_templates shapes."""
    lib = load()
    nds = np.ascontiguousarray(nds, np.float64)
    els0 = np.ascontiguousarray(els0, np.int32)
    mat = np.ascontiguousarray(mat, np.int32).ravel()
    mult = np.ascontiguousarray(mult, np.int64).ravel()
    n_el = els0.shape[0]
    el_tmpl = np.empty(n_el, np.int32)
    first = np.empty(max(int(max_templates), 1), np.int64)
    jf = np.empty(n_el, np.float64) if with_jfactor else None
    nt = C.c_int32()
    st = lib.bfe_element_templates(int(device), nds.shape[0], nds.shape[1], ptr(nds, C.c_double),
                                   n_el, els0.shape[1], ptr(els0, C.c_int32), ptr(mat, C.c_int32),
                                   ptr(mult, C.c_int64), int(max_templates),
                                   ptr(el_tmpl, C.c_int32), ptr(first, C.c_int64), C.byref(nt),
                                   ptr(jf, C.c_double))
    if st != BFE_OK:
        raise BfeError(st, lib.bfe_assemble_last_error().decode())
    if nt.value == 0:
        return None
    return el_tmpl, first[:nt.value].copy(), jf


def convolve(f, g, g_per_signal, device=0):
    """bfe_convolve: f (n_sig, n_t) row-major; g (n_g,) or (n_sig, n_g)."""
    lib = load()
    f = np.ascontiguousarray(f, np.float64)
    g = np.ascontiguousarray(g, np.float64)
    n_sig, n_t = f.shape
    n_g = g.shape[-1] if g.size else 0
    h = np.empty((n_sig, n_t), np.float64)
    check(lib.bfe_convolve(int(device), n_sig, n_t, ptr(f, C.c_double), n_g,
                           ptr(g, C.c_double) if g.size else None, 1 if g_per_signal else 0,
                           ptr(h, C.c_double)))
    return h


def disps_to_forces(K, Cm, m, dt, disp, lyrs, in_or_out, solver_mode, device=0,
                    no_alloc=False):
    """bfe_disps_to_forces with scipy CSR K, C; disp (n, n_t). Returns
    (force_idx 0-based, forces (n_force, n_t)). no_alloc=True calls
    bfe_disps_to_forces_noalloc (no device allocation; at most 64 MiB of
    device arrays)."""
    lib = load()
    fn = lib.bfe_disps_to_forces_noalloc if no_alloc else lib.bfe_disps_to_forces
    n = K.shape[0]
    K = K.tocsr()
    Cm = Cm.tocsr()
    kp, kc, kv = (np.ascontiguousarray(K.indptr, np.int64), np.ascontiguousarray(K.indices, np.int32),
                  np.ascontiguousarray(K.data, np.float64))
    cp, cc, cv = (np.ascontiguousarray(Cm.indptr, np.int64),
                  np.ascontiguousarray(Cm.indices, np.int32),
                  np.ascontiguousarray(Cm.data, np.float64))
    m = np.ascontiguousarray(m, np.float64)
    disp = np.asfortranarray(disp, np.float64)
    lyrs = np.ascontiguousarray(lyrs, np.int32).ravel()
    n_t = disp.shape[1]
    nf = int(np.count_nonzero((lyrs == 2) | (lyrs == 3)))
    idx = np.empty(nf, np.int32)
    out = np.empty((nf, n_t), np.float64, order="F")
    check(fn(int(device), n, ptr(kp, C.c_int64), ptr(kc, C.c_int32), ptr(kv, C.c_double),
             ptr(cp, C.c_int64), ptr(cc, C.c_int32), ptr(cv, C.c_double), ptr(m, C.c_double),
             float(dt), n_t, disp.ctypes.data_as(_f64p), ptr(lyrs, C.c_int32), int(in_or_out),
             int(solver_mode), nf, ptr(idx, C.c_int32), out.ctypes.data_as(_f64p)))
    return idx, out


def assemble_device(nds, els0, el_type, el_mat0, el_abs, mat_D, mat_rho, dof_mask, p_law,
                    max_damping, msr, device=0):
    """Calls bfe_assemble_device/fill (GPU). 0-based element nodes and
    materials (ignored for ASI2D2). Returns (k_rowptr, k_col, k_val, c_rowptr,
    c_col, c_val, m, gl_lookup)."""
    lib = load()
    nds = np.ascontiguousarray(nds, np.float64)
    els0 = np.ascontiguousarray(els0, np.int32)
    el_type = np.ascontiguousarray(el_type, np.int32)
    el_mat0 = np.ascontiguousarray(el_mat0, np.int32)
    el_abs = np.ascontiguousarray(el_abs, np.float64)
    mat_D = np.ascontiguousarray(mat_D, np.float64)
    mat_rho = np.ascontiguousarray(mat_rho, np.float64)
    n_dof, zk, zc, max_df, work = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32(), C.c_void_p()
    st = lib.bfe_assemble_device(
        int(device), nds.shape[0], nds.shape[1], ptr(nds, C.c_double), els0.shape[0],
        els0.shape[1], ptr(els0, C.c_int32), ptr(el_type, C.c_int32), ptr(el_mat0, C.c_int32),
        ptr(el_abs, C.c_double), mat_rho.size, ptr(mat_D, C.c_double), ptr(mat_rho, C.c_double),
        int(dof_mask), float(p_law), float(max_damping), float(msr), C.byref(n_dof), C.byref(zk),
        C.byref(zc), C.byref(max_df), C.byref(work))
    if st != BFE_OK:
        raise BfeError(st, lib.bfe_assemble_last_error().decode())
    n = n_dof.value
    out = (np.empty(n + 1, np.int64), np.empty(zk.value, np.int32), np.empty(zk.value, np.float64),
           np.empty(n + 1, np.int64), np.empty(zc.value, np.int32), np.empty(zc.value, np.float64),
           np.empty(n, np.float64), np.empty((nds.shape[0], max_df.value), np.int64))
    st = lib.bfe_assemble_device_fill(
        work, ptr(out[0], C.c_int64), ptr(out[1], C.c_int32), ptr(out[2], C.c_double),
        ptr(out[3], C.c_int64), ptr(out[4], C.c_int32), ptr(out[5], C.c_double),
        ptr(out[6], C.c_double), ptr(out[7], C.c_int64))
    if st != BFE_OK:
        raise BfeError(st, lib.bfe_assemble_last_error().decode())
    return out

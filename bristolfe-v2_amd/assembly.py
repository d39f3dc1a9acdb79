"""Global matrix assembly (FP64 host set-up).

Mirror of code/internal/fn_build_global_matrices_v5.m:1-167 with the same
inputs, options and outputs, vectorised and chunked so that multi-million
element meshes assemble in seconds:

  [K, C, M, gl_lookup] = fn_build_global_matrices_v5(nds, els, el_mat_i,
                                                     el_abs_i, el_typ_i, matls, fe_options)

Conventions kept from the reference:
  * node and DOF numbers are 1-based in ``els`` and ``gl_lookup``
    (0 in ``gl_lookup`` = absent), global matrices are 0-based scipy CSR;
  * unique element types and materials are processed in sorted order
    (:51, :75); el_mat_i == 0 means "no material" (interface elements, :92-95);
  * absorbing layers: C_e += M_e * a^p * max_damping and
    K_e *= exp(ln(msr) * a^(p+1)) (:115-116);
  * global index (node-1)*max_df + df (:127-128), duplicates summed (:145-147);
  * DOFs whose column is empty in K, M and C are removed, keeping order
    (fn_reduce_global_matrices.m:7-16); gl_lookup maps (node, dof) to the
    compacted index (fn_create_fast_lookup.m:4-8);
  * optional interface damping (fn_add_interface_damping.m:18-53).

Element mass matrices are diagonal for every element type here, so M (and
the diagonal part of C) is accumulated as a vector; only K and off-diagonal
C go through coordinate assembly. Exact zeros are not stored, as in MATLAB.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .elements import element_function
from .options import set_default_fields

_CHUNK = 1 << 20  # elements per vectorised batch

DEFAULTS = {
    "dof_to_use": [],
    "damping_power_law": 3,
    "max_damping": 3.1415e7,
    "max_stiffness_reduction": 0.01,
    "interface_damping_factor": 0,
    # "numpy": vectorised coordinate assembly (any element type);
    # "native": libbfe_hip's row-by-row host assembler (CPE3/AC2D3/C3D4, no
    # interface elements); "device": the same on the GPU, with ASI2D2
    # interface elements too; "auto": the GPU for more than 200k elements
    # when one is present, else native for more than 2M elements without
    # interface elements, else numpy.
    "assembly_backend": "auto",
}

_NATIVE_TYPES = {"CPE3": 0, "AC2D3": 1, "C3D4": 2}
_DEVICE_TYPES = {"CPE3": 0, "AC2D3": 1, "C3D4": 2, "ASI2D2": 3}


def find_dof_in_use_and_max_dof_per_el(unique_typs, dof_to_use=None):
    """Mirror of code/internal/fn_find_dof_in_use_and_max_dof_per_el.m:1-21."""
    dofs = []
    max_per_el = 0
    for t in unique_typs:
        _, loc_df = element_function(t)(None, None, None, None, dof_to_use)
        dofs.extend(int(d) for d in loc_df)
        max_per_el = max(max_per_el, len(loc_df))
    return np.unique(np.asarray(dofs, dtype=np.int64)), max_per_el


def _material(matls, mi):
    if mi <= 0:
        return 0.0, 0.0
    mat = matls[mi - 1]
    D = np.asarray(mat["D"], dtype=np.float64)
    rho = float(mat["density"] if "density" in mat else mat["rho"])
    if D.shape == (3, 3):  # legacy 3x3 plane-strain D (:99-102)
        D6 = np.zeros((6, 6))
        idx = [0, 1, 5]
        D6[np.ix_(idx, idx)] = D
        D = D6
    return D, rho


def fn_build_global_matrices_v5(nds, els, el_mat_i, el_abs_i, el_typ_i, matls, fe_options=None):
    fe_options = set_default_fields(dict(fe_options or {}), DEFAULTS)
    nds = np.asarray(nds, dtype=np.float64)
    els = np.asarray(els, dtype=np.int64)
    el_mat_i = np.asarray(el_mat_i, dtype=np.int64).ravel()
    el_abs_i = np.asarray(el_abs_i, dtype=np.float64).ravel()
    el_typ_i = np.asarray(el_typ_i, dtype=object).ravel()
    if el_mat_i.size != els.shape[0]:
        raise ValueError("Length of element materials vector must equal number of elements")
    if el_abs_i.size != els.shape[0]:
        raise ValueError("Length of absorbing index vector must equal number of elements")

    dof_to_use = fe_options["dof_to_use"]
    unique_typs, typ_code = _type_codes(el_typ_i)
    backend = fe_options["assembly_backend"]
    device_ok = set(unique_typs) <= set(_DEVICE_TYPES)
    if backend == "auto" and device_ok and els.shape[0] > 200_000:
        try:
            return _assemble_device_with_options(nds, els, el_mat_i, el_abs_i, el_typ_i,
                                                 unique_typs, typ_code, matls, fe_options)
        except _no_gpu_errors() as e:  # no GPU (or no library): assemble on the host
            if getattr(e, "status", None) not in (None, _hip_status()):
                raise
    native_ok = set(unique_typs) <= set(_NATIVE_TYPES) and not fe_options["interface_damping_factor"]
    if backend == "native" and not native_ok:
        raise ValueError("native assembly supports CPE3, AC2D3 and C3D4 elements only")
    if backend == "native" or (backend == "auto" and native_ok and els.shape[0] > 2_000_000):
        return _assemble_native(nds, els, el_mat_i, el_abs_i, unique_typs, typ_code, matls, fe_options)
    if backend == "device":
        if not device_ok:
            raise ValueError("device assembly supports CPE3, AC2D3, C3D4 and ASI2D2 elements only")
        return _assemble_device_with_options(nds, els, el_mat_i, el_abs_i, el_typ_i, unique_typs,
                                             typ_code, matls, fe_options)
    unique_df, _ = find_dof_in_use_and_max_dof_per_el(unique_typs, dof_to_use)
    max_df = int(unique_df.max())
    no_nds = nds.shape[0]
    total_dof = no_nds * max_df
    p_law = float(fe_options["damping_power_law"])
    max_damp = float(fe_options["max_damping"])
    log_msr = np.log(float(fe_options["max_stiffness_reduction"]))

    m_diag = np.zeros(total_dof)
    c_diag = np.zeros(total_dof)
    Ki, Kj, Kv, Ci, Cj, Cv = [], [], [], [], [], []
    for ti, typ in enumerate(unique_typs):
        fn = element_function(typ)
        el_i = typ_code == ti
        for mi in np.unique(el_mat_i[el_i]):
            sel = np.nonzero((el_mat_i == mi) & el_i)[0]
            if sel.size == 0:
                continue
            D, rho = _material(matls, int(mi))
            for s0 in range(0, sel.size, _CHUNK):
                s = sel[s0:s0 + _CHUNK]
                nn = {"CPE3": 3, "AC2D3": 3, "ASI2D2": 2, "C3D4": 4}[typ]
                e = els[s, :nn] - 1
                el_K, el_C, el_M, loc_nd, loc_df = fn(nds, e, D, rho, dof_to_use)
                a = el_abs_i[s]
                if np.any(a != 0):
                    el_C = el_C + el_M * (a ** p_law * max_damp)[:, None, None]
                    el_K = el_K * np.exp(log_msr * a ** (p_law + 1))[:, None, None]
                gi = e[:, np.asarray(loc_nd) - 1] * max_df + (np.asarray(loc_df) - 1)  # (m, k)
                kk = len(loc_nd)
                d = np.arange(kk)
                # diagonal element mass / damping
                m_diag += np.bincount(gi.ravel(), weights=el_M[:, d, d].ravel(), minlength=total_dof)
                c_diag += np.bincount(gi.ravel(), weights=el_C[:, d, d].ravel(), minlength=total_dof)
                off = ~np.eye(kk, dtype=bool)
                if np.any(el_M[:, off]):
                    raise NotImplementedError("non-diagonal element mass matrices")
                rows = np.broadcast_to(gi[:, :, None], (gi.shape[0], kk, kk))
                cols = np.broadcast_to(gi[:, None, :], (gi.shape[0], kk, kk))
                if np.any(el_K):
                    nzk = el_K != 0
                    Ki.append(rows[nzk].astype(np.int64))
                    Kj.append(cols[nzk].astype(np.int64))
                    Kv.append(el_K[nzk])
                offC = el_C * off[None]
                if np.any(offC):
                    nzc = offC != 0
                    Ci.append(rows[nzc].astype(np.int64))
                    Cj.append(cols[nzc].astype(np.int64))
                    Cv.append(offC[nzc])

    def _coo(I, J, V):
        if not I:
            return sp.csr_matrix((total_dof, total_dof))
        A = sp.coo_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))),
                          shape=(total_dof, total_dof)).tocsr()
        A.sum_duplicates()
        A.eliminate_zeros()
        return A

    K = _coo(Ki, Kj, Kv)
    del Ki, Kj, Kv
    C = _coo(Ci, Cj, Cv) + sp.diags(c_diag, format="csr")
    C.eliminate_zeros()
    M = sp.diags(m_diag, format="csr")
    M.eliminate_zeros()

    # fn_reduce_global_matrices: keep DOFs with a non-empty column
    valid = np.zeros(total_dof, dtype=bool)
    for A in (K, M, C):
        valid[A.indices[A.data != 0]] = True
    K, M, C = _reduce(K, valid), _reduce(M, valid), _reduce(C, valid)

    gl_nds = np.repeat(np.arange(1, no_nds + 1), max_df)[valid]
    gl_dofs = np.tile(np.arange(1, max_df + 1), no_nds)[valid]
    gl_lookup = fn_create_fast_lookup(gl_nds, gl_dofs, no_nds, 0)

    if fe_options["interface_damping_factor"]:
        C = fn_add_interface_damping(els, el_typ_i, K, C, M, gl_lookup,
                                     fe_options["interface_damping_factor"])
    return K, C, M, gl_lookup


def _type_codes(el_typ_i):
    """Sorted distinct element types and each element's index into them. A
    mesh of one element type (the common case on large meshes) is detected
    with one vectorised comparison instead of a Python pass per element."""
    if el_typ_i.size and bool(np.all(el_typ_i == el_typ_i[0])):
        return [el_typ_i[0]], np.zeros(el_typ_i.size, dtype=np.int32)
    uniq = sorted(set(el_typ_i.tolist()))
    lut = {t: i for i, t in enumerate(uniq)}
    return uniq, np.fromiter((lut[t] for t in el_typ_i), dtype=np.int32, count=el_typ_i.size)


def _assemble_native(nds, els, el_mat_i, el_abs_i, unique_typs, typ_code, matls, fe_options):
    from . import _lib

    if el_mat_i.size and el_mat_i.min() < 1:
        raise ValueError("native assembly needs a material for every element")
    D = np.zeros((len(matls), 36))
    rho = np.zeros(len(matls))
    for i, m in enumerate(matls):
        Dm, r = _material(matls, i + 1)
        Dm = np.asarray(Dm, dtype=np.float64)
        D[i, :Dm.size] = Dm.ravel() if Dm.ndim else [float(Dm)]
        rho[i] = r
    mask = 0
    for d in (fe_options["dof_to_use"] or []):
        mask |= 1 << (int(d) - 1)
    typ = np.asarray([_NATIVE_TYPES[t] for t in unique_typs], dtype=np.int32)[typ_code]
    nn = max(3, els.shape[1])
    els0 = np.full((els.shape[0], nn), -1, dtype=np.int32)
    for s0 in range(0, els.shape[0], _CHUNK):
        e = els[s0:s0 + _CHUNK]
        els0[s0:s0 + e.shape[0], :els.shape[1]] = np.where(e > 0, e - 1, -1)
    rowptr, col, kval, m, c, gl = _lib.assemble_native(
        nds, els0, typ, el_mat_i - 1, el_abs_i, D, rho, mask, fe_options["damping_power_law"],
        fe_options["max_damping"], fe_options["max_stiffness_reduction"])
    n = m.size
    K = sp.csr_matrix((kval, col, rowptr), shape=(n, n))
    K.has_sorted_indices = True
    M = sp.diags(m, format="csr")
    C = sp.diags(c, format="csr")
    C.eliminate_zeros()
    return K, C, M, gl


def _material_tables(matls):
    D = np.zeros((max(len(matls), 1), 36))
    rho = np.zeros(max(len(matls), 1))
    for i, m in enumerate(matls):
        Dm, r = _material(matls, i + 1)
        Dm = np.asarray(Dm, dtype=np.float64)
        D[i, :Dm.size] = Dm.ravel() if Dm.ndim else [float(Dm)]
        rho[i] = r
    return D, rho


def _els0(els):
    nn = max(3, els.shape[1])
    els0 = np.full((els.shape[0], nn), -1, dtype=np.int32)
    for s0 in range(0, els.shape[0], _CHUNK):
        e = els[s0:s0 + _CHUNK]
        els0[s0:s0 + e.shape[0], :els.shape[1]] = np.where(e > 0, e - 1, -1)
    return els0


def _no_gpu_errors():
    from . import _lib
    return (_lib.BfeError, OSError, RuntimeError)


def _hip_status():
    from . import _lib
    return _lib.BFE_ERR_HIP


def _assemble_device_with_options(nds, els, el_mat_i, el_abs_i, el_typ_i, unique_typs, typ_code,
                                  matls, fe_options):
    K, C, M, gl = _assemble_device(nds, els, el_mat_i, el_abs_i, unique_typs, typ_code, matls,
                                   fe_options)
    if fe_options["interface_damping_factor"]:
        C = fn_add_interface_damping(els, el_typ_i, K, C, M, gl,
                                     fe_options["interface_damping_factor"])
    return K, C, M, gl


def _assemble_device(nds, els, el_mat_i, el_abs_i, unique_typs, typ_code, matls, fe_options):
    """GPU assembly (bfe_assemble_device): K and C as CSR, M diagonal."""
    from . import _lib

    typ = np.asarray([_DEVICE_TYPES[t] for t in unique_typs], dtype=np.int32)[typ_code]
    solid = typ != _DEVICE_TYPES["ASI2D2"]
    if np.any(el_mat_i[solid] < 1):
        raise ValueError("device assembly needs a material for every non-interface element")
    D, rho = _material_tables(matls)
    mask = 0
    for d in (fe_options["dof_to_use"] or []):
        mask |= 1 << (int(d) - 1)
    kr, kc, kv, cr, cc, cv, m, gl = _lib.assemble_device(
        nds, _els0(els), typ, np.where(solid, el_mat_i - 1, -1), el_abs_i, D, rho, mask,
        fe_options["damping_power_law"], fe_options["max_damping"],
        fe_options["max_stiffness_reduction"])
    n = m.size
    K = sp.csr_matrix((kv, kc, kr), shape=(n, n))
    K.has_sorted_indices = True
    C = sp.csr_matrix((cv, cc, cr), shape=(n, n))
    C.has_sorted_indices = True
    M = sp.diags(m, format="csr")
    M.eliminate_zeros()
    return K, C, M, gl


def _reduce(A, valid):
    new = np.cumsum(valid) - 1
    A = A[np.nonzero(valid)[0]]  # row selection
    A = sp.csr_matrix((A.data, new[A.indices], A.indptr), shape=(A.shape[0], A.shape[0]))
    A.has_sorted_indices = True
    return A


def fn_create_fast_lookup(global_matrix_nodes, global_matrix_dofs, max_nds, max_dofs):
    """Mirror of code/internal/fn_create_fast_lookup.m:1-8 (1-based contents)."""
    gn = np.asarray(global_matrix_nodes, dtype=np.int64)
    gd = np.asarray(global_matrix_dofs, dtype=np.int64)
    nr = max(int(gn.max()) if gn.size else 0, int(max_nds))
    nc = max(int(gd.max()) if gd.size else 0, int(max_dofs))
    lut = np.zeros((nr, nc), dtype=np.int64)
    lut[gn - 1, gd - 1] = np.arange(1, gn.size + 1)
    return lut


def fn_add_interface_damping(els, el_typ_i, K, C, M, gl_lookup, interface_damping_factor=1.0):
    """Mirror of code/internal/fn_add_interface_damping.m:18-53."""
    dK = K.diagonal()
    dM = M.diagonal()
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.sqrt(dK / dM)
    level = np.nanmin(ratio) / np.pi ** 2 * interface_damping_factor
    is_int = np.asarray([t == "ASI2D2" for t in np.asarray(el_typ_i, dtype=object).ravel()])
    nds_int = np.asarray(els)[is_int]
    nds_int = nds_int[nds_int > 0]
    if nds_int.size == 0:
        return C
    gi = gl_lookup[nds_int - 1, :].ravel()
    gi = np.unique(gi[gi > 0]) - 1
    add = np.zeros(C.shape[0])
    add[gi] = dM[gi] * level
    C = (C + sp.diags(add, format="csr")).tocsr()
    C.eliminate_zeros()
    return C

"""Element library, vectorised over elements (FP64 host set-up).

Each ``el_<TYPE>(nds, els, D, rho, dofs_to_use)`` returns
``(el_K, el_C, el_M, loc_nd, loc_df)`` with matrices of shape
(n_els, k, k); ``loc_nd``/``loc_df`` are 1-based local node and DOF numbers.
Calling with ``nds=None`` returns only ``(loc_nd, loc_df)`` after the DOF
filter, as the reference's element functions do for empty arguments.

  CPE3   code/internal/elements/fn_el_CPE3.m:26-158   plane-strain triangle
  AC2D3  code/internal/elements/fn_el_AC2D3.m:26-80   acoustic triangle
  ASI2D2 code/internal/elements/fn_el_ASI2D2.m:26-82  fluid-solid interface
  C3D4   (no counterpart in the reference) constant-strain tetrahedron

Closed forms used (J = signed 2 x area, fn_el_CPE3.m:56):
  CPE3:  Ke = (J/2) B^T D B, B in engineering Voigt order
         (e11, e22, e33, g23, g13, g12) with d/dz = 0
         (el_builder/el_sym_CPE3.m:58-64); lumped M = J rho / 6 per DOF.
  AC2D3: Ke_ij = -(b_i b_j + c_i c_j) / (2 J rho); M_ii = -J / (6 K).
  ASI2D2: C(u1_a, p_b) = (y2 - y1)/4, C(u2_a, p_b) = (x1 - x2)/4 (symmetric).
  C3D4:  Ke = (J/6) B^T D B with J = signed 6 x volume, lumped M = J rho / 24.
"""
from __future__ import annotations

import numpy as np

ELEMENT_TYPES = ("AC2D3", "ASI2D2", "C3D4", "CPE3")


def _filter_dofs(loc_nd, loc_df, dofs_to_use, mats=()):
    loc_nd = np.asarray(loc_nd)
    loc_df = np.asarray(loc_df)
    if dofs_to_use is None or len(dofs_to_use) == 0:
        dofs_to_use = np.unique(loc_df)
    keep = np.isin(loc_df, np.asarray(dofs_to_use))
    out = [m[:, keep][:, :, keep] for m in mats]
    return loc_nd[keep], loc_df[keep], out


def _tri_geometry(nds, els):
    x = nds[:, 0][els]  # (m, 3)
    y = nds[:, 1][els]
    x1, x2, x3 = x[:, 0], x[:, 1], x[:, 2]
    y1, y2, y3 = y[:, 0], y[:, 1], y[:, 2]
    J = x1 * y2 - y1 * x2 - x1 * y3 + y1 * x3 + x2 * y3 - y2 * x3
    b = np.stack([y2 - y3, y3 - y1, y1 - y2], axis=1)
    c = np.stack([x3 - x2, x1 - x3, x2 - x1], axis=1)
    return J, b, c


def el_CPE3(nds, els, D, rho, dofs_to_use=None):
    loc_nd = [1, 1, 1, 2, 2, 2, 3, 3, 3]
    loc_df = [1, 2, 3, 1, 2, 3, 1, 2, 3]
    ln, ld, _ = _filter_dofs(loc_nd, loc_df, dofs_to_use)
    if nds is None or els is None or len(els) == 0:
        return ln, ld
    keep = np.nonzero(np.isin(loc_df, ld))[0]
    D = np.asarray(D, dtype=np.float64)
    J, b, c = _tri_geometry(nds, els)
    m = els.shape[0]
    B = np.zeros((m, 6, 9))
    for i in range(3):
        bi = b[:, i] / J
        ci = c[:, i] / J
        B[:, 0, 3 * i + 0] = bi          # e11 = du1/dx
        B[:, 1, 3 * i + 1] = ci          # e22 = du2/dy
        B[:, 3, 3 * i + 2] = ci          # g23 = du3/dy
        B[:, 4, 3 * i + 2] = bi          # g13 = du3/dx
        B[:, 5, 3 * i + 0] = ci          # g12 = du1/dy + du2/dx
        B[:, 5, 3 * i + 1] = bi
    B = np.ascontiguousarray(B[:, :, keep])
    rows = np.nonzero(np.abs(B).sum(axis=(0, 2)) > 0)[0]   # strain rows that are used
    B = np.ascontiguousarray(B[:, rows, :])
    Dr = D[np.ix_(rows, rows)]
    DB = np.matmul(Dr[None], B)
    K = np.matmul(B.transpose(0, 2, 1), DB)
    K *= (J / 2.0)[:, None, None]
    k = keep.size
    M = np.zeros((m, k, k))
    idx = np.arange(k)
    M[:, idx, idx] = (J * rho / 6.0)[:, None]
    return K, np.zeros_like(M), M, ln, ld


def el_AC2D3(nds, els, D, rho, dofs_to_use=None):
    loc_nd = [1, 2, 3]
    loc_df = [4, 4, 4]
    if nds is None or els is None or len(els) == 0:
        ln, ld, _ = _filter_dofs(loc_nd, loc_df, dofs_to_use)
        return ln, ld
    bulk = float(np.asarray(D).ravel()[0])
    J, b, c = _tri_geometry(nds, els)
    K = -(b[:, :, None] * b[:, None, :] + c[:, :, None] * c[:, None, :]) / (2.0 * J * rho)[:, None, None]
    m = els.shape[0]
    M = np.zeros((m, 3, 3))
    M[:, [0, 1, 2], [0, 1, 2]] = (-J / (6.0 * bulk))[:, None]
    Cm = np.zeros_like(M)
    ln, ld, (K, Cm, M) = _filter_dofs(loc_nd, loc_df, dofs_to_use, (K, Cm, M))
    return K, Cm, M, ln, ld


def el_ASI2D2(nds, els, D, rho, dofs_to_use=None):
    loc_nd = [1, 1, 1, 1, 2, 2, 2, 2]
    loc_df = [1, 2, 3, 4, 1, 2, 3, 4]
    if nds is None or els is None or len(els) == 0:
        ln, ld, _ = _filter_dofs(loc_nd, loc_df, dofs_to_use)
        return ln, ld
    x1, y1 = nds[els[:, 0], 0], nds[els[:, 0], 1]
    x2, y2 = nds[els[:, 1], 0], nds[els[:, 1], 1]
    m = els.shape[0]
    Cm = np.zeros((m, 8, 8))
    cx = y2 / 4.0 - y1 / 4.0   # couples u1 of either node with p of either node
    cy = x1 / 4.0 - x2 / 4.0   # couples u2 with p
    for a in (0, 4):           # ux, uy of node a at local a, a+1
        for pb in (3, 7):      # p of node b
            Cm[:, a, pb] = cx
            Cm[:, pb, a] = cx
            Cm[:, a + 1, pb] = cy
            Cm[:, pb, a + 1] = cy
    K = np.zeros_like(Cm)
    M = np.zeros_like(Cm)
    ln, ld, (K, Cm, M) = _filter_dofs(loc_nd, loc_df, dofs_to_use, (K, Cm, M))
    return K, Cm, M, ln, ld


def el_C3D4(nds, els, D, rho, dofs_to_use=None):
    loc_nd = [1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4]
    loc_df = [1, 2, 3] * 4
    if nds is None or els is None or len(els) == 0:
        ln, ld, _ = _filter_dofs(loc_nd, loc_df, dofs_to_use)
        return ln, ld
    D = np.asarray(D, dtype=np.float64)
    X = nds[els]  # (m, 4, 3)
    E = X[:, 1:, :] - X[:, :1, :]  # rows: x2-x1, x3-x1, x4-x1
    J = np.linalg.det(E)           # 6 x signed volume
    # grad N_i = inverse(E) applied to reference gradients
    Einv = np.linalg.inv(E)        # (m, 3, 3): columns give d(xi)/dx
    gref = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    G = np.einsum("ik,mjk->mij", gref, Einv)  # (m, 4, 3): dN_i/dx_j
    m = els.shape[0]
    B = np.zeros((m, 6, 12))
    for i in range(4):
        gx, gy, gz = G[:, i, 0], G[:, i, 1], G[:, i, 2]
        B[:, 0, 3 * i + 0] = gx
        B[:, 1, 3 * i + 1] = gy
        B[:, 2, 3 * i + 2] = gz
        B[:, 3, 3 * i + 1] = gz          # g23 = du2/dz + du3/dy
        B[:, 3, 3 * i + 2] = gy
        B[:, 4, 3 * i + 0] = gz          # g13 = du1/dz + du3/dx
        B[:, 4, 3 * i + 2] = gx
        B[:, 5, 3 * i + 0] = gy          # g12 = du1/dy + du2/dx
        B[:, 5, 3 * i + 1] = gx
    DB = np.einsum("ab,mbj->maj", D, B)
    K = np.einsum("mai,maj->mij", B, DB) * (J / 6.0)[:, None, None]
    M = np.zeros((m, 12, 12))
    idx = np.arange(12)
    M[:, idx, idx] = (J * rho / 24.0)[:, None]
    Cm = np.zeros_like(M)
    ln, ld, (K, Cm, M) = _filter_dofs(loc_nd, loc_df, dofs_to_use, (K, Cm, M))
    return K, Cm, M, ln, ld


ELEMENTS = {"CPE3": el_CPE3, "AC2D3": el_AC2D3, "ASI2D2": el_ASI2D2, "C3D4": el_C3D4}


def element_function(name: str):
    try:
        return ELEMENTS[name]
    except KeyError:
        raise ValueError(f"unknown element type {name!r}") from None

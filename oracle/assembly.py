"""Per-element loop restatement of fn_build_global_matrices_v5.m.

ORACLE — test infrastructure only (see oracle/__init__.py). For small meshes.

  :51-54    DOF set of the element types (after dof_to_use), max_df
  :70-106   loop over sorted types, sorted materials; element matrices
  :115-116  C_e += M_e a^p max_damping; K_e *= exp(ln(msr) a^(p+1))
  :127-128  global index (node - 1) * max_df + dof
  :145-147  duplicates summed (dictionary accumulation)
  fn_reduce_global_matrices.m:7-16   keep DOFs with a non-empty column
  fn_create_fast_lookup.m:4-8        gl_lookup
"""
from __future__ import annotations

from collections import defaultdict

import numpy as np
import scipy.sparse as sp

from .elements import NODES_PER_EL, element


def _dofs_of(typ, dof_to_use):
    dummy = {"CPE3": [1, 2, 3], "AC2D3": [4], "ASI2D2": [1, 2, 3, 4], "C3D4": [1, 2, 3]}[typ]
    if dof_to_use is None or len(dof_to_use) == 0:
        return dummy
    return [d for d in dummy if d in dof_to_use]


def build_global_matrices(nds, els, el_mat_i, el_abs_i, el_typ_i, matls, fe_options=None):
    o = {"dof_to_use": [], "damping_power_law": 3, "max_damping": 3.1415e7,
         "max_stiffness_reduction": 0.01}
    o.update(fe_options or {})
    dof_to_use = o["dof_to_use"]
    typs = sorted(set(el_typ_i))
    dofs = sorted({d for t in typs for d in _dofs_of(t, dof_to_use)})
    max_df = max(dofs)
    n_nodes = len(nds)
    total = n_nodes * max_df
    K, C, M = defaultdict(float), defaultdict(float), defaultdict(float)
    for t in typs:
        mats_here = sorted({int(el_mat_i[e]) for e in range(len(els)) if el_typ_i[e] == t})
        for mi in mats_here:
            if mi > 0:
                mat = matls[mi - 1]
                D = np.asarray(mat["D"], dtype=np.float64)
                rho = mat["rho"]
            else:
                D, rho = 0.0, 0.0
            for e in range(len(els)):
                if el_typ_i[e] != t or int(el_mat_i[e]) != mi:
                    continue
                nodes = [int(v) for v in els[e][:NODES_PER_EL[t]]]
                coords = [tuple(nds[v - 1]) for v in nodes]
                Ke, Ce, Me, ln, ld = element(t, coords, D, rho)
                keep = [k for k in range(len(ld)) if (not dof_to_use) or ld[k] in dof_to_use]
                a = float(el_abs_i[e])
                damp = a ** o["damping_power_law"] * o["max_damping"]
                red = np.exp(np.log(o["max_stiffness_reduction"]) * a ** (o["damping_power_law"] + 1))
                for i in keep:
                    gi = (nodes[ln[i] - 1] - 1) * max_df + ld[i] - 1
                    for j in keep:
                        gj = (nodes[ln[j] - 1] - 1) * max_df + ld[j] - 1
                        if Ke[i, j] != 0:
                            K[gi, gj] += Ke[i, j] * red
                        c = Ce[i, j] + Me[i, j] * damp
                        if c != 0:
                            C[gi, gj] += c
                        if Me[i, j] != 0:
                            M[gi, gj] += Me[i, j]

    def to_csr(dct):
        if not dct:
            return sp.csr_matrix((total, total))
        keys = np.array(list(dct.keys()), dtype=np.int64)
        vals = np.array(list(dct.values()))
        A = sp.csr_matrix((vals, (keys[:, 0], keys[:, 1])), shape=(total, total))
        A.eliminate_zeros()
        return A

    Kg, Cg, Mg = to_csr(K), to_csr(C), to_csr(M)
    colsum = sum(np.asarray(abs(A).sum(axis=0)).ravel() for A in (Kg, Mg, Cg))
    valid = colsum > 0
    idx = np.nonzero(valid)[0]
    Kg, Cg, Mg = (A[idx][:, idx].tocsr() for A in (Kg, Cg, Mg))
    gl = np.zeros((n_nodes, max_df), dtype=np.int64)
    for new, g in enumerate(idx):
        gl[g // max_df, g % max_df] = new + 1
    return Kg, Cg, Mg, gl


def add_interface_damping(els, el_typ_i, K, C, M, gl_lookup, factor=1.0):
    """code/internal/fn_add_interface_damping.m:18-53 on dense arrays, line by
    line: level = min(sqrt(diag K ./ diag M)) / pi^2 * factor (:30, NaN
    skipped as MATLAB's min does), gi = the interface nodes' global indices
    in MATLAB's column-major order, duplicates kept (:33-45), and
    C(gi, gi) = C(gi, gi) + M(gi, gi) * level (:52), whose duplicate
    positions are assigned in order, the last one kept (numpy's fancy
    assignment does the same). Returns a dense C."""
    Kd, Cd, Md = (np.asarray(sp.csr_matrix(A).toarray()) for A in (K, C, M))
    with np.errstate(divide="ignore", invalid="ignore"):
        level = np.nanmin(np.sqrt(np.diag(Kd) / np.diag(Md))) / np.pi ** 2 * factor
    typ = np.asarray(el_typ_i, dtype=object).ravel()
    nds = np.asarray(els)[typ == "ASI2D2", :]
    nds = nds.ravel(order="F")
    nds = nds[nds > 0]
    if nds.size == 0:
        return Cd
    gi = np.asarray(gl_lookup)[nds - 1, :].ravel(order="F")
    gi = gi[gi > 0] - 1
    ix = np.ix_(gi, gi)
    Cd[ix] = Cd[ix] + Md[ix] * level
    return Cd

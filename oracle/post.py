"""Field output and stability restatements.

ORACLE — test infrastructure only (see oracle/__init__.py).

field_output_ke     code/internal/fn_get_field_output.m:64-79, :109-135
    KE_g = |v^2 diag(M)| / 2 per global DOF, summed over the DOFs of a node,
    averaged over the nodes of an element, then sqrt.
max_amplification   tests/fn_amplification_matrix.m:1-42
    spectral radius of the 2n x 2n one-step amplification matrix.
"""
from __future__ import annotations

import numpy as np


def field_output_ke(f_out, els, diag_m, gl_lookup):
    f_out = np.asarray(f_out, dtype=np.float64)
    n_nodes, n_dof = gl_lookup.shape
    KE_n = np.zeros((n_nodes, f_out.shape[1]))
    for node in range(n_nodes):
        for d in range(n_dof):
            g = gl_lookup[node, d]
            if g > 0:
                KE_n[node] += np.abs(f_out[g - 1] ** 2 * diag_m[g - 1]) / 2
    out = np.zeros((len(els), f_out.shape[1]))
    for e, row in enumerate(np.asarray(els)):
        nodes = [v for v in row if v > 0]
        out[e] = sum(KE_n[v - 1] for v in nodes) / len(nodes)
    return np.sqrt(out)


def max_amplification(M, C, K, dt, solver_mode):
    M, C, K = M.toarray(), C.toarray(), K.toarray()
    n = K.shape[0]
    iM = np.diag(1.0 / M.sum(axis=0))
    I = np.eye(n)
    if solver_mode == "vel at last half time step":
        B1 = 2 * I - dt * iM @ C - dt ** 2 * iM @ K
        B2 = dt * iM @ C - I
        A = np.block([[B1, B2], [I, np.zeros((n, n))]])
    else:
        B1 = 2 * I - dt ** 2 * iM @ K
        B2 = -I + dt / 2 * iM @ C
        B3 = I + dt / 2 * iM @ C
        A = np.block([[np.linalg.solve(B3, B1), np.linalg.solve(B3, B2)], [I, np.zeros((n, n))]])
    return float(np.max(np.abs(np.linalg.eigvals(A))))

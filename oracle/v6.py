"""Literal restatement of code/internal/fn_explicit_dynamic_solver_v6.m.

ORACLE — test infrastructure only (see oracle/__init__.py).

Line map:
  v6:93        dt = time(2) - time(1)
  v6:112-120   field frames at ti = 1:N:n_t, zero-initialised
  v6:122-123   diag_M = spdiags(sum(M)), inv_M = spdiags(1 ./ sum(M))
  v6:130-139   A, B, D per solver mode; 'X \\ Y' for the current-step mode
  v6:163-166   ti_start = min(find(sum(abs(F)))): with several rows the first
               column holding a non-zero entry; a single row sums to one
               scalar, so ti_start = 1 whenever any sample is non-zero
  v6:180-213   f(idx) = F(:, ti); u = A*f + B*u1 + D*u2; hist(:, ti) = u(h);
               frame (u - u1)/dt or u; shift
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

FIELD_U_TYPES = {"mean(u1)", "mean(u2)", "mean(u3)", "curl(u)", "div(u)", "raw(u)"}


def left_divide(X, Y):
    """X \\ Y for sparse X that is diagonal outside a (small) coupled block.

    Block Gaussian elimination: rows/columns with off-diagonal entries of X
    form the block I, solved densely with LU after scaling each row by the
    sum of its absolute values; every other row is divided by its diagonal.
    The same linear algebra as MATLAB's sparse '\\' (UMFPACK, whose default
    row scaling is that sum). The scaling matters on the fluid-solid
    interface, where |m| spans 13 orders of magnitude: X_II's condition
    number is about 1e10 unscaled and 3e5 scaled, and an unscaled LU is off
    by 1e-12 per solve.
    """
    X = sp.csr_matrix(X)
    Y = sp.csr_matrix(Y)
    n = X.shape[0]
    coo = X.tocoo()
    nz_off = (coo.row != coo.col) & (coo.data != 0)
    coupled = np.zeros(n, dtype=bool)
    coupled[coo.row[nz_off]] = True
    coupled[coo.col[nz_off]] = True
    d = X.diagonal()
    scale = np.zeros(n)
    scale[~coupled] = 1.0 / d[~coupled]
    out = (sp.diags(scale) @ Y).tocoo()
    rows, cols, vals = [out.row], [out.col], [out.data]
    I = np.nonzero(coupled)[0]
    if I.size:
        XII = X[I][:, I].toarray()
        YI = Y[I].toarray()
        r = 1.0 / np.abs(XII).sum(axis=1)
        Z = sla.lu_solve(sla.lu_factor(r[:, None] * XII), r[:, None] * YI)
        zr, zc = np.nonzero(Z)
        rows.append(I[zr])
        cols.append(zc)
        vals.append(Z[zr, zc])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, Y.shape[1]))


def abd_matrices(K, C, M, dt, solver_mode):
    ndf = K.shape[0]
    m = np.asarray(M.sum(axis=0)).ravel()
    inv_M = sp.diags(1.0 / m)
    I = sp.eye(ndf, format="csr")
    if solver_mode == "vel at last half time step":
        A = dt ** 2 * inv_M
        B = 2 * I - dt * (inv_M @ C) - dt ** 2 * (inv_M @ K)
        D = dt * (inv_M @ C) - I
    elif solver_mode in ("vel at curent time step", "vel at current time step"):
        X = I + dt / 2 * (inv_M @ C)
        A = left_divide(X, dt ** 2 * inv_M)
        B = left_divide(X, 2 * I - dt ** 2 * (inv_M @ K))
        D = left_divide(X, -I + dt / 2 * (inv_M @ C))
    else:
        raise ValueError(solver_mode)
    return sp.csr_matrix(A), sp.csr_matrix(B), sp.csr_matrix(D)


def abd_diagonal_x(K, C, M, dt, solver_mode, chunk=1 << 20):
    """A, B, D of v6:130-139 for a diagonal C (so X = I + dt/2 M^-1 C is
    diagonal and X \\ Y divides row i by X_ii), formed row block by row
    block on K's sparsity pattern: the same matrices as abd_matrices, for
    models too large for its generic sparse products (the 24M-DOF cube has
    9e8 entries). Every row of K must hold its diagonal entry."""
    K = sp.csr_matrix(K)
    n = K.shape[0]
    Cd = sp.csr_matrix(C)
    c = Cd.diagonal()
    if (Cd - sp.diags(c)).count_nonzero():
        raise ValueError("abd_diagonal_x needs a diagonal C")
    m = np.asarray(M.sum(axis=0)).ravel()
    inv_m = 1.0 / m
    cur = solver_mode in ("vel at curent time step", "vel at current time step")
    if not cur and solver_mode != "vel at last half time step":
        raise ValueError(solver_mode)
    x = 1.0 + dt / 2 * (inv_m * c) if cur else np.ones(n)
    kd = dt ** 2 * inv_m                      # rows of dt^2 M^-1 K
    bd = 2.0 if cur else 2.0 - dt * (inv_m * c)   # the diagonal of 2I (- dt M^-1 C)
    data = np.empty_like(K.data)
    ip = K.indptr
    for r0 in range(0, n, chunk):
        r1 = min(n, r0 + chunk)
        p0, p1 = ip[r0], ip[r1]
        rows = np.repeat(np.arange(r0, r1), np.diff(ip[r0:r1 + 1]))
        d = -(kd[rows] * K.data[p0:p1])
        on = K.indices[p0:p1] == rows
        if np.count_nonzero(on) != r1 - r0:
            raise ValueError("every row of K must hold its diagonal entry")
        d[on] += bd if np.isscalar(bd) else bd[rows[on]]
        data[p0:p1] = d / x[rows]
    B = sp.csr_matrix((data, K.indices, K.indptr), shape=K.shape)
    if cur:
        A = sp.diags(dt ** 2 * inv_m / x, format="csr")
        D = sp.diags((-1.0 + dt / 2 * (inv_m * c)) / x, format="csr")
    else:
        A = sp.diags(dt ** 2 * inv_m, format="csr")
        D = sp.diags(dt * (inv_m * c) - 1.0, format="csr")
    return A, B, D


def ti_start_of(forcing_functions):
    """v6:163-166, ``min(find(sum(abs(forcing_functions))))``, 1-based; None
    when find is empty. MATLAB's ``sum`` runs down the first non-singleton
    dimension: down each column of a p x n_t matrix (p > 1), but along the
    row of a 1 x n_t signal, which gives one scalar and so ti_start = 1 as
    soon as any sample is non-zero (NaN included)."""
    F = np.atleast_2d(np.asarray(forcing_functions, dtype=np.float64))
    if F.shape[0] == 1:
        return 1 if np.any(F != 0) else None
    nz = np.nonzero(np.abs(F).sum(axis=0) != 0)[0]
    return int(nz[0]) + 1 if nz.size else None


def explicit_dynamic_solver_v6(K, C, M, time, forcing_indices, forcing_functions,
                               history_indices, field_output_every_n_frames=np.inf,
                               field_output_type="KE", solver_mode="vel at curent time step",
                               abd=None, t_end=None):
    """1-based indices as in MATLAB. Returns (hist, fld, fld_time, ti_start).
    t_end (1-based, inclusive) stops the loop early: a timing sample of the
    first steps; later history columns stay 0."""
    time = np.asarray(time, dtype=np.float64).ravel()
    n_t = time.size
    ndf = K.shape[0]
    dt = time[1] - time[0]
    fi = np.asarray(forcing_indices, dtype=np.int64).ravel() - 1
    F = np.atleast_2d(np.asarray(forcing_functions, dtype=np.float64))
    hi = np.asarray(history_indices, dtype=np.int64).ravel() - 1
    hist = np.zeros((hi.size, n_t))
    if np.isinf(field_output_every_n_frames):
        frames = np.zeros(0, dtype=np.int64)
    else:
        frames = np.arange(1, n_t + 1, int(field_output_every_n_frames))
    fld = np.zeros((ndf, frames.size))
    fld_time = np.zeros(frames.size)
    is_vel = field_output_type not in FIELD_U_TYPES
    A, B, D = abd if abd is not None else abd_matrices(K, C, M, dt, solver_mode)
    u1 = np.zeros(ndf)
    u2 = np.zeros(ndf)
    f = np.zeros(ndf)
    ti_start = ti_start_of(F) if fi.size else None
    if ti_start is None:
        return hist, fld, fld_time, n_t + 1
    frame_of = {int(t): k for k, t in enumerate(frames)}
    for ti in range(ti_start, (n_t if t_end is None else min(n_t, t_end)) + 1):
        col = F[:, ti - 1]
        if col.size == 1:
            f[fi] = col[0]
        else:
            for k in range(fi.size):  # sequential assignment: last duplicate wins
                f[fi[k]] = col[k]
        u = A @ f + B @ u1 + D @ u2
        if hi.size:
            hist[:, ti - 1] = u[hi]
        if ti in frame_of:
            k = frame_of[ti]
            fld_time[k] = time[ti - 1]
            fld[:, k] = (u - u1) / dt if is_vel else u
        u2 = u1
        u1 = u
    return hist, fld, fld_time, ti_start

"""Python driver of the oracle's C march (oracle/march.c).

ORACLE — test infrastructure only (see oracle/__init__.py).

Per-row coefficients follow v6:130-139 with the diagonal of
X = I + dt/2 M^-1 C eliminated (current-step mode):
  non-interface rows: alpha = dt^2/(m + dt c/2), beta = 2m/(m + dt c/2),
                      gamma = -(m - dt c/2)/(m + dt c/2)
  interface rows:     alpha = dt^2/m, beta = 2, gamma = -1 + dt c/(2m), then
                      u_I = X_II^-1 (w_I + dt/2 M^-1 Coff u2)
last-half-step mode:  alpha = dt^2/m, beta = 2 - dt c/m, gamma = dt c/m - 1,
                      u_I = w_I - dt M^-1 Coff (u1 - u2)
X_II^-1 comes from a SciPy sparse LU (splu) of X_II, all columns, entries
below 1e-18 of a column's largest dropped.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .v6 import ti_start_of

LIB = Path(__file__).resolve().parent / "liboracle_march.so"
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from . import build_march
            build_march.build()
        _lib = C.CDLL(str(LIB))
        _lib.oracle_march.restype = C.c_int
    return _lib


class OracleMarch:
    def __init__(self, K, C_mat, M, dt, solver_mode="vel at curent time step"):
        K = sp.csr_matrix(K)
        K.sort_indices()
        n = K.shape[0]
        m = np.asarray(M.sum(axis=0)).ravel()
        Cc = sp.csr_matrix(C_mat) if C_mat is not None else sp.csr_matrix((n, n))
        c = Cc.diagonal()
        Coff = (Cc - sp.diags(c)).tocsr()
        Coff.eliminate_zeros()
        coupled = np.zeros(n, dtype=bool)
        cc = Coff.tocoo()
        coupled[cc.row] = True
        coupled[cc.col] = True
        I = np.nonzero(coupled)[0]
        mode0 = solver_mode != "vel at last half time step"
        if mode0:
            den = m + 0.5 * dt * c
            alpha = np.where(coupled, dt * dt / m, dt * dt / den)
            beta = np.where(coupled, 2.0, 2.0 * m / den)
            gamma = np.where(coupled, -1.0 + 0.5 * dt * c / m, -(m - 0.5 * dt * c) / den)
            sq, sr = 0.0, 0.5 * dt
        else:
            alpha = dt * dt / m
            beta = 2.0 - dt * c / m
            gamma = dt * c / m - 1.0
            sq, sr = -dt, dt
        self.n, self.dt = n, dt
        self.K = K
        self.alpha, self.beta, self.gamma = alpha, beta, gamma
        self.I = I.astype(np.int32)
        self.sq, self.sr = sq, sr
        Cs = (sp.diags(1.0 / m) @ Coff).tocsr()[I] if I.size else sp.csr_matrix((0, n))
        self.Cs = sp.csr_matrix(Cs)
        self.X = None
        if I.size and mode0:
            X = (sp.eye(n) + 0.5 * dt * sp.diags(1.0 / m) @ Cc).tocsr()[I][:, I].tocsc()
            lu = spla.splu(X)
            rows, cols, vals = [], [], []
            for c0 in range(0, I.size, 512):
                c1 = min(c0 + 512, I.size)
                E = np.zeros((I.size, c1 - c0))
                E[np.arange(c0, c1), np.arange(c1 - c0)] = 1.0
                Z = lu.solve(E)
                mx = np.abs(Z).max(axis=0)
                keep = np.abs(Z) > 1e-18 * mx[None, :]
                r, cidx = np.nonzero(keep)
                rows.append(r)
                cols.append(cidx + c0)
                vals.append(Z[r, cidx])
            self.X = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                   shape=(I.size, I.size))

    def run(self, n_t, frc_idx0, F, hist_idx0, t_end=None, threads=None):
        """0-based indices. Returns (hist q x n_t, ti_start 1-based)."""
        lib = _load()
        if threads is not None:
            os.environ["OMP_NUM_THREADS"] = str(threads)
        F = np.atleast_2d(np.asarray(F, dtype=np.float64))
        fidx = np.ascontiguousarray(frc_idx0, np.int32).ravel()
        hidx = np.ascontiguousarray(hist_idx0, np.int32).ravel()
        ts = ti_start_of(F) if fidx.size else None
        q = hidx.size
        hist = np.zeros((q, n_t), order="F")
        if ts is None:
            return hist, n_t + 1
        t0 = ts - 1
        Fc = np.asfortranarray(F)
        K = self.K
        d = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        kptr = K.indptr.astype(np.int64)
        kcol = K.indices.astype(np.int32)
        kval = np.ascontiguousarray(K.data, np.float64)
        cptr = self.Cs.indptr.astype(np.int64)
        ccol = self.Cs.indices.astype(np.int32)
        cval = np.ascontiguousarray(self.Cs.data, np.float64)
        if self.X is not None:
            xptr = self.X.indptr.astype(np.int64)
            xcol = self.X.indices.astype(np.int32)
            xval = np.ascontiguousarray(self.X.data, np.float64)
            xa = (d(xptr, C.c_int64), d(xcol, C.c_int32), d(xval, C.c_double))
        else:
            xa = (None, None, None)
        rc = lib.oracle_march(
            C.c_int64(self.n), d(kptr, C.c_int64), d(kcol, C.c_int32), d(kval, C.c_double),
            d(self.alpha, C.c_double), d(self.beta, C.c_double), d(self.gamma, C.c_double),
            C.c_int32(self.I.size), d(self.I, C.c_int32), d(cptr, C.c_int64), d(ccol, C.c_int32),
            d(cval, C.c_double), C.c_double(self.sq), C.c_double(self.sr), *xa,
            C.c_int32(n_t), C.c_int32(t0), C.c_int32(n_t if t_end is None else t_end),
            C.c_int32(fidx.size), d(fidx, C.c_int32), d(Fc, C.c_double), C.c_int32(F.shape[0]),
            C.c_int32(q), d(hidx, C.c_int32), d(hist, C.c_double))
        if rc != 0:
            raise MemoryError("oracle march allocation failed")
        return hist, t0 + 1

"""Subdomain coupling restatements.

ORACLE — test infrastructure only (see oracle/__init__.py).

convolve          code/internal/fn_convolve.m:22-34
    FFTs of length 2 n (n = size(f, dim)) of f and g, their product, the
    inverse FFT, the first n samples; the real part is returned.
disps_to_forces   code/fn_convert_disps_to_forces_v2.m:1-29
    force set = layers 2 and 3; displacements of layers 3, 4 ('in') or
    1, 2 ('out') zeroed; accn, vel and the one-step-delayed disp by
    differences of [0, 0, disp_inp]; M(set,:) accn + C(set,:) vel +
    K(set,:) disp with M = diag(sum(M)).
parse_to_fmc      subdoms/fn_run_main_model.m:239-256
    for each (tx, rx) pair the monitored rows of the transmitter's run that
    belong to the receiver's nodes, summed (or weighted by wt), convolved
    with the input signal when one is given.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def convolve(f, g, dim):
    f = np.atleast_2d(np.asarray(f, dtype=np.float64))
    g = np.asarray(g, dtype=np.float64)
    if g.ndim == 1:  # one signal along dim
        g = g[None, :] if dim == 2 else g[:, None]
    ax = dim - 1
    P = 2 * f.shape[ax]
    F = np.fft.fft(f, P, axis=ax)
    G = np.fft.fft(g, P, axis=ax)
    h = np.fft.ifft(F * G, P, axis=ax)
    return np.real(h)[: f.shape[0], : f.shape[1]]


def disps_to_forces(K_sub, C_sub, M_sub, time_step, disp_inp, lyrs, in_or_out, solver_mode):
    lyrs = np.asarray(lyrs).ravel()
    force_in_set = lyrs == 2
    force_out_set = lyrs == 3
    force_set = force_in_set | force_out_set
    if in_or_out == "in":
        freeze_set = (lyrs == 3) | (lyrs == 4)
    elif in_or_out == "out":
        freeze_set = (lyrs == 1) | (lyrs == 2)
    else:
        raise ValueError(in_or_out)
    disp_inp = np.array(disp_inp, dtype=np.float64)
    disp_inp[freeze_set, :] = 0
    n = disp_inp.shape[0]
    M_sub = sp.diags(np.asarray(M_sub.sum(axis=0)).ravel(), 0, shape=(n, n), format="csr")
    tmp = np.hstack([np.zeros((n, 2)), disp_inp])
    accn = (tmp[:, 2:] - 2 * tmp[:, 1:-1] + tmp[:, :-2]) / time_step ** 2
    if solver_mode == "vel at last half time step":
        vel = (tmp[:, 1:-1] - tmp[:, :-2]) / time_step
    elif solver_mode == "vel at curent time step":
        vel = (tmp[:, 2:] - tmp[:, :-2]) / (2 * time_step)
    else:
        raise ValueError(solver_mode)
    disp = np.hstack([np.zeros((n, 1)), disp_inp[:, :-1]])
    K_sub, C_sub = sp.csr_matrix(K_sub), sp.csr_matrix(C_sub)
    force_inp = M_sub[force_set] @ accn + C_sub[force_set] @ vel + K_sub[force_set] @ disp
    return force_inp, force_set, force_in_set, force_out_set


def parse_to_fmc(tx, rx, time_pts, steps, res, trans, in_sig):
    """tx, rx: 1-based transducer numbers of each FMC pair (the template)."""
    time_data = np.zeros((time_pts, len(tx)))
    for k in range(len(tx)):
        t, r = tx[k] - 1, rx[k] - 1
        i = np.isin(np.asarray(steps[t]["mon"]["nds"]), np.asarray(trans[r]["nds"]))
        tmp = np.asarray(res[t]["dsps"])[i, :]
        if "wt" in trans[r]:
            tmp = np.asarray(trans[r]["wt"]).ravel() @ tmp
        else:
            tmp = tmp.sum(axis=0)
        time_data[:, k] = tmp
        if in_sig is not None and len(in_sig):
            time_data[:, k] = convolve(time_data[:, k][:, None],
                                       np.asarray(in_sig, dtype=np.float64)[:, None], 1)[:, 0]
    return time_data

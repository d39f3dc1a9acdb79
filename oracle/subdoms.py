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


def _mappings(dom_mod, main_res):
    # subdoms/fn_run_subdomain_model.m:119-145
    lyrs = np.asarray(dom_mod["bdry_lyrs"])
    bdry_nds = np.nonzero(lyrs > 0)[0] + 1
    nb_m = np.asarray(dom_mod["main_nd_i"])[bdry_nds - 1]
    mon_nds = np.asarray(main_res["mon_nds"])
    mn_res_i = np.nonzero(np.isin(mon_nds, nb_m))[0]
    gl_i = np.asarray(main_res["dsp_gi"])[mn_res_i]
    Nb_m = mon_nds[mn_res_i]
    near = np.abs(Nb_m[:, None] - nb_m[None, :]).argmin(axis=1)  # interp1(..., 'nearest')
    b_nds = bdry_nds[near]
    return mn_res_i, gl_i, b_nds, np.asarray(main_res["mon_dfs"])[mn_res_i], lyrs[b_nds - 1]


def run_subdomain_model(main, fe_options, build_global_matrices, march):
    """subdoms/fn_run_subdomain_model.m:26-116, line by line, for checking:
    per subdomain and transmitter the incident boundary displacements
    convolved with the input signal, 'in' forces, one march of the subdomain
    model per transmitter (fn_FE_entry_point -> fn_BristolFE_v2 -> v6), 'out'
    forces and, per receiver, the sum of the receiver's impulse responses at
    the force DOFs convolved with the forces, added to the pristine FMC.
    build_global_matrices(mod, matls, opts) -> (K, C, M, gl_lookup) and
    march(K, C, M, time, frc_gi, frcs, hist_gi) -> hist (1-based indices) are
    the oracle's. Returns the FMC time_data of every subdomain."""
    o = dict(fe_options)
    tx, rx, doms = o["tx_trans"], o["rx_trans"], o["doms_to_run"]
    time = np.asarray(main["inp"]["time"], dtype=np.float64)
    sig = np.asarray(main["inp"]["sig"], dtype=np.float64).ravel()
    dt = time[1] - time[0]
    res_m = main["res"]
    K, C, M = res_m["mats"]["K"], res_m["mats"]["C"], res_m["mats"]["M"]
    mn_all = np.stack([np.asarray(res_m["mon_nds"]), np.asarray(res_m["mon_dfs"])], axis=1)
    out = []
    for d in doms:
        dm = main["doms"][d - 1]["mod"]
        Kd, Cd, Md, gl = build_global_matrices(dm, main["matls"], o)
        fmc = np.array(res_m["fmc"]["time_data"], copy=True)
        dsps_all, frce_sets = [], []
        mn_res_i, gl_i, b_nds, b_dfs, b_lyrs = _mappings(dm, res_m)
        g0 = gl_i - 1
        K_sub, C_sub, M_sub = K[g0][:, g0], C[g0][:, g0], M[g0][:, g0]
        for t in tx:
            bd = np.asarray(res_m["trans"][t - 1]["dsps"])[mn_res_i]
            bd = convolve(bd, sig, 2)
            frcs, frce_set, _, _ = disps_to_forces(K_sub, C_sub, M_sub, dt, bd, b_lyrs, "in",
                                                   o["solver_mode"])
            # fn_BristolFE_v2.m:76-84, 159-168: (node, dof) -> global index, invalid dropped
            fgi = gl[b_nds[frce_set] - 1, b_dfs[frce_set] - 1]
            fvalid = fgi > 0
            hgi = gl[b_nds - 1, b_dfs - 1]
            hvalid = hgi > 0
            h = march(Kd, Cd, Md, time, fgi[fvalid], frcs[fvalid], hgi[hvalid])
            dsps = np.zeros((b_nds.size, time.size))
            dsps[hvalid] = h
            dsps_all.append(dsps)
            frce_sets.append(frce_set)
        for si, t in enumerate(tx):
            frcs, frce_set, _, _ = disps_to_forces(K_sub, C_sub, M_sub, dt, dsps_all[si], b_lyrs,
                                                   "out", o["solver_mode"])
            mn_nds_i = np.asarray(dm["main_nd_i"])[b_nds[frce_set] - 1]
            want = {(int(a), int(b)) for a, b in zip(mn_nds_i, b_dfs[frce_set])}
            i = np.asarray([(int(a), int(b)) in want for a, b in mn_all], dtype=bool)
            for r in rx:
                tmp = convolve(np.asarray(res_m["trans"][r - 1]["dsps"])[i], frcs, 2).sum(axis=0)
                k = np.nonzero((np.asarray(res_m["fmc"]["tx"]) == t) &
                               (np.asarray(res_m["fmc"]["rx"]) == r))[0]
                fmc[:, k] += tmp[:, None]
        out.append(fmc)
    return out

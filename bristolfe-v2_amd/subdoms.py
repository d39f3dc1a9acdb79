"""Subdomain coupling workflow: the operations around the batched march.

  fn_convolve(f, g, dim)                          code/internal/fn_convolve.m:1-40
  fn_convert_disps_to_forces_v2(K_sub, C_sub, M_sub, time_step, disp_inp,
                                lyrs, in_or_out, solver_mode)
                                                  code/fn_convert_disps_to_forces_v2.m:1-29
  fn_convert_to_step_data(...)                    subdoms/fn_run_main_model.m:259-267
  fn_create_fmc_template(main, fe_options)        subdoms/fn_run_main_model.m:187-236
  fn_parse_to_fmc(fmc_template, steps, res, trans, in_sig, fe_options)
                                                  subdoms/fn_run_main_model.m:239-256
  fn_run_main_model(main, fe_options)             subdoms/fn_run_main_model.m:1-131,
                                                  'impulse response' mode: the
                                                  pristine FMC and the
                                                  displacements at the subdomain
                                                  boundaries

The convolution and the force conversion run on the GPU (bfe_convolve,
bfe_disps_to_forces); the main model's transmitters are marched together by
bfe_run_batched, 64 load cases per pass, instead of one march per
transmitter. ``main`` is a dict with the reference's fields: ``mod``
(nds, els, el_mat_i, el_abs_i, el_typ_i), ``matls``, ``trans`` (list of
{"nds", "dfs", ["wt"]}, 1-based), optional ``doms`` (list of {"mod":
{"main_nd_i", "bdry_lyrs"}}) and ``inp`` (time, sig) once the model has run.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib
from .assembly import find_dof_in_use_and_max_dof_per_el, fn_build_global_matrices_v5
from .mesh import fn_gaussian_pulse
from .options import set_default_fields
from .solver import make_plan

_MODES = {"vel at curent time step": _lib.MODE_VEL_AT_CURRENT,
          "vel at last half time step": _lib.MODE_VEL_AT_LAST_HALF}


def fn_convolve(f, g, dim, use_gpu_if_present=1):
    """f convolved with g along MATLAB dimension ``dim`` (1: down the columns,
    2: along the rows) through FFTs of length 2 size(f, dim), the first
    size(f, dim) samples kept. g is one signal (a vector) or one per signal
    (the shape of f). Always on the GPU (use_gpu_if_present is accepted for
    the reference's signature); the result is real."""
    f = np.asarray(f, dtype=np.float64)
    squeeze = f.ndim == 1
    if squeeze:
        f = f[:, None] if dim == 1 else f[None, :]
    if dim not in (1, 2) or f.ndim != 2:
        raise ValueError("fn_convolve: f must be a vector or matrix and dim 1 or 2")
    g = np.asarray(g, dtype=np.float64)
    rows = f if dim == 2 else f.T                    # signals as rows, time contiguous
    if g.ndim == 1 or (1 in g.shape and g.size != rows.size):
        g_rows, per = g.ravel(), False
    else:
        g_rows = g if dim == 2 else g.T
        if g_rows.shape[0] != rows.shape[0]:
            raise ValueError("fn_convolve: g must be one signal or one per signal of f")
        per = True
    h = _lib.convolve(rows, g_rows, per)
    h = h if dim == 2 else h.T
    return h.ravel() if squeeze else h


def fn_convert_disps_to_forces_v2(K_sub, C_sub, M_sub, time_step, disp_inp, lyrs, in_or_out,
                                  solver_mode):
    """Forces on the DOFs of layers 2 and 3 that reproduce disp_inp (n x n_t)
    in the subdomain's matrices, with layers 3, 4 ('in') or 1, 2 ('out')
    held at zero. Returns (force_inp, force_set, force_in_set, force_out_set)
    as the reference does (boolean masks over the n DOFs)."""
    if in_or_out not in ("in", "out"):
        raise ValueError("in_or_out must be 'in' or 'out'")
    if solver_mode not in _MODES:
        raise ValueError(f"unknown solver mode {solver_mode!r}")
    lyrs = np.asarray(lyrs).ravel()
    m = np.asarray(sp.csr_matrix(M_sub).sum(axis=0)).ravel()   # M = spdiags(sum(M)) (:13)
    _, force = _lib.disps_to_forces(sp.csr_matrix(K_sub), sp.csr_matrix(C_sub), m, time_step,
                                    disp_inp, lyrs, 0 if in_or_out == "in" else 1,
                                    _MODES[solver_mode])
    force_in_set = lyrs == 2
    force_out_set = lyrs == 3
    return np.ascontiguousarray(force), force_in_set | force_out_set, force_in_set, force_out_set


def fn_convert_to_step_data(t, inp, frc_nds, frc_dfs, mon_nds, mon_dfs, f_every):
    return {"load": {"time": t, "frcs": inp, "frc_nds": frc_nds, "frc_dfs": frc_dfs},
            "mon": {"nds": mon_nds, "dfs": mon_dfs, "field_output_every_n_frames": f_every}}


def fn_create_fmc_template(main, fe_options):
    """tx, rx of every pair (meshgrid order), zero time data and the array
    geometry (element centres and extrema)."""
    tx, rx = np.meshgrid(np.asarray(fe_options["tx_trans"]), np.asarray(fe_options["rx_trans"]))
    tmpl = {"tx": tx.ravel(order="F"), "rx": rx.ravel(order="F"),
            "time_data": np.zeros((int(fe_options["time_pts"]), tx.size)),
            "time": np.asarray(main["inp"]["time"]).ravel()}
    nds = np.asarray(main["mod"]["nds"])
    ne = len(main["trans"])
    arr = {"centre_freq": fe_options["centre_freq"], "el_type": "rectangular"}
    for k in ("el_xc", "el_yc", "el_zc", "el_x1", "el_y1", "el_z1", "el_x2", "el_y2", "el_z2"):
        arr[k] = np.zeros(ne)
    for t, tr in enumerate(main["trans"]):
        p = nds[np.asarray(tr["nds"]).ravel() - 1]
        c, mx = p.mean(axis=0), p.max(axis=0)
        if c.size == 2:
            arr["el_xc"][t], arr["el_zc"][t] = c
            arr["el_x1"][t], arr["el_z1"][t] = mx
            arr["el_x2"][t], arr["el_y2"][t], arr["el_z2"][t] = c[0], 1.0, c[1]
        else:
            arr["el_xc"][t], arr["el_yc"][t], arr["el_zc"][t] = c
            arr["el_x1"][t], arr["el_y1"][t], arr["el_z1"][t] = mx[0], c[1], c[2]
            arr["el_x2"][t], arr["el_y2"][t], arr["el_z2"][t] = c[0], mx[1], c[2]
    tmpl["array"] = arr
    return tmpl


def fn_parse_to_fmc(fmc_template, steps, res, trans, in_sig, fe_options=None):
    """Each (tx, rx) pair's A-scan: the transmitter run's monitored rows on the
    receiver's nodes, summed or weighted by the receiver's wt, then (pristine
    impulse responses) convolved with the input signal, all pairs in one
    batched convolution."""
    fmc = dict(fmc_template)
    tx, rx = np.asarray(fmc["tx"]), np.asarray(fmc["rx"])
    n_t = fmc["time_data"].shape[0]
    A = np.zeros((tx.size, n_t))
    for k in range(tx.size):
        t, r = int(tx[k]) - 1, int(rx[k]) - 1
        i = np.isin(np.asarray(steps[t]["mon"]["nds"]), np.asarray(trans[r]["nds"]))
        tmp = np.asarray(res[t]["dsps"])[i, :]
        A[k] = np.asarray(trans[r]["wt"]).ravel() @ tmp if "wt" in trans[r] else tmp.sum(axis=0)
    if in_sig is not None and np.size(in_sig):
        A = fn_convolve(A, np.asarray(in_sig, dtype=np.float64).ravel(), 2)
    fmc["time_data"] = np.ascontiguousarray(A.T)
    return fmc


def fn_run_main_model(main, fe_options=None):
    """Pristine model of the subdomain method in impulse-response mode: every
    transmitter fires an impulse, the transducer nodes and the subdomain
    boundary nodes are monitored, and the FMC is the receivers' responses
    convolved with the input signal. Returns main with inp, res (fmc,
    trans[e].dsps at the monitored DOFs, mon_nds, mon_dfs, dsp_gi, mats)."""
    mod = dict(main["mod"])
    if "el_typ_i" not in mod:
        mod["el_typ_i"] = np.asarray([main["matls"][i - 1]["el_typ"] for i in mod["el_mat_i"]],
                                     dtype=object)
    ne = len(main["trans"])
    defaults = {"centre_freq": mod.get("design_centre_freq"), "number_of_cycles": 5,
                "time_step": mod.get("max_safe_time_step"), "time_pts": 1000, "max_time": None,
                "dof_to_use": [], "tx_trans": list(range(1, ne + 1)),
                "rx_trans": list(range(1, ne + 1)), "max_damping": None,
                "damping_power_law": 3, "max_stiffness_reduction": 0.01,
                "solver_mode": "vel at curent time step", "solver_precision": "double"}
    o = set_default_fields(dict(fe_options or {}), defaults)
    if o["centre_freq"] is None or o["time_step"] is None:
        raise ValueError("centre_freq and time_step are needed (fe_options or mod)")
    if o["max_damping"] is None:
        o["max_damping"] = o["centre_freq"] * 2 * np.pi
    o["dof_to_use"] = find_dof_in_use_and_max_dof_per_el(
        sorted(set(np.asarray(mod["el_typ_i"]).tolist())), o["dof_to_use"])[0].tolist()
    if o["max_time"]:
        o["time_pts"] = int(np.ceil(o["max_time"] / o["time_step"]))
    n_t = int(o["time_pts"])
    time = np.arange(n_t) * o["time_step"]
    main = dict(main)
    main["inp"] = {"time": time, "sig": fn_gaussian_pulse(time, o["centre_freq"],
                                                         o["number_of_cycles"])}
    K, C, M, gl = fn_build_global_matrices_v5(mod["nds"], mod["els"], mod["el_mat_i"],
                                              mod.get("el_abs_i", np.zeros(len(mod["els"]))),
                                              mod["el_typ_i"], main["matls"], o)
    # monitored nodes: subdomain boundaries and transducers (:58-73)
    flag = np.zeros(np.asarray(mod["nds"]).shape[0], dtype=bool)
    for d in main.get("doms", []):
        dm = d["mod"]
        flag[np.asarray(dm["main_nd_i"])[np.asarray(dm["bdry_lyrs"]) > 0] - 1] = True
    for tr in main["trans"]:
        flag[np.asarray(tr["nds"]).ravel() - 1] = True
    mon_nds = np.nonzero(flag)[0] + 1
    mn, md = np.meshgrid(mon_nds, np.asarray(o["dof_to_use"]))
    mon_nds, mon_dfs = mn.ravel(order="F"), md.ravel(order="F")
    gi = np.zeros(mon_nds.size, dtype=np.int64)
    ok = (mon_dfs <= gl.shape[1])
    gi[ok] = gl[mon_nds[ok] - 1, mon_dfs[ok] - 1]
    valid = gi > 0
    hist = gi[valid] - 1
    impulse = np.zeros(n_t)
    impulse[0] = 1.0
    steps = []
    frc_dof, frc_src = [], []
    for e, tr in enumerate(main["trans"]):
        steps.append(fn_convert_to_step_data(time, impulse, tr["nds"], tr["dfs"], mon_nds,
                                             mon_dfs, np.inf))
        fn_ = np.asarray(tr["nds"]).ravel()
        fd = np.broadcast_to(np.asarray(tr["dfs"]).ravel(), fn_.shape)
        g = gl[fn_ - 1, fd - 1]
        if np.any(g <= 0):
            raise ValueError(f"transducer {e + 1} forces a DOF the model does not have")
        frc_dof.append(g - 1)
        frc_src.append(np.full(g.size, e))
    frc_dof, frc_src = np.concatenate(frc_dof), np.concatenate(frc_src)
    plan = make_plan(K, C, M, time[1] - time[0], o["solver_mode"], o["solver_precision"])
    try:
        dsps = np.zeros((hist.size, ne, n_t))
        for c0 in range(0, ne, 64):           # 64 load cases per batched march
            c1 = min(c0 + 64, ne)
            sel = (frc_src >= c0) & (frc_src < c1)
            h, _ = plan.run_batched(n_t, c1 - c0, frc_dof[sel], frc_src[sel] - c0, impulse, hist)
            dsps[:, c0:c1, :] = h
    finally:
        plan.close()
    # the reference's dsps rows are the valid monitors (fn_parse_to_bdry_nds, :167-176)
    fe_res = [{"dsps": dsps[:, e, :], "valid_mon_dsps": valid, "dsp_gi": gi[valid]}
              for e in range(ne)]
    res = {"mats": {"K": K, "C": C, "M": M, "gl_lookup": gl}, "dsp_gi": gi[valid],
           "mon_nds": mon_nds[valid], "mon_dfs": mon_dfs[valid],
           "trans": [{"dsps": r["dsps"]} for r in fe_res]}
    steps_valid = [dict(s, mon={"nds": mon_nds[valid], "dfs": mon_dfs[valid]}) for s in steps]
    tmpl = fn_create_fmc_template(main, o)
    res["fmc_template"] = tmpl
    res["fmc"] = fn_parse_to_fmc(tmpl, steps_valid, fe_res, main["trans"], main["inp"]["sig"], o)
    main["res"] = res
    return main

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
  fn_create_subdomain(mn_mod, matls, inner_bdry, abs_layer_thick)
                                                  subdoms/fn_create_subdomain.m:1-87
  fn_run_subdomain_model(main, fe_options)        subdoms/fn_run_subdomain_model.m:1-116:
                                                  every transmitter's incident
                                                  field injected at each
                                                  subdomain's boundary layers, all
                                                  transmitters marched together,
                                                  and the scattered field added
                                                  to the FMC through the
                                                  receivers' impulse responses

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
from .mesh import (fn_add_fluid_solid_interface_els, fn_calc_element_centres,
                   fn_dist_point_to_bdry_2D, fn_elements_in_region, fn_find_free_edges,
                   fn_gaussian_pulse, fn_remove_unused_nodes)
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


def _adjacent_els(els, consider, choose):
    """(adjacent, common nodes) of subdoms/fn_create_subdomain.m:93-108: the
    elements of `choose` that share a node with an element of `consider`, and
    those shared nodes as MATLAB's els(tmp) lists them (column-major order,
    repeats kept)."""
    un = np.unique(els[consider])
    un = un[un != 0]
    tmp = np.isin(els, un) & choose[:, None]
    common = els.ravel(order="F")[tmp.ravel(order="F")]
    return tmp.any(axis=1), common


def fn_create_subdomain(mn_mod, matls, inner_bdry, abs_layer_thick, abs_layer_start_bdry=None):
    """Subdomain model of the main model's elements inside inner_bdry
    (subdoms/fn_create_subdomain.m:1-87): four rings of elements around them
    give the boundary layers 1-4 of their nodes (bdry_lyrs), the nodes of
    layer 3 start the absorbing region (unless abs_layer_start_bdry is
    given), whose elements carry el_abs_i = distance / abs_layer_thick and
    end at 1; the interface elements are made again on the new mesh.
    Returns the model with nds, els, el_mat_i, el_abs_i, el_typ_i,
    bdry_lyrs, main_nd_i (1-based main node of each node), main_int_el_i,
    inner_bndry_pts, outer_bndry_pts and int_el_i."""
    nds = np.asarray(mn_mod["nds"], dtype=np.float64)
    els = np.asarray(mn_mod["els"])
    mat = np.asarray(mn_mod["el_mat_i"])
    if "el_typ_i" in mn_mod:
        typ = np.asarray(mn_mod["el_typ_i"], dtype=object)
    else:
        typ = np.asarray([matls[i - 1]["el_typ"] for i in mat], dtype=object)
    bdry_lyrs = np.zeros(nds.shape[0], dtype=np.int64)
    el_abs = np.zeros(els.shape[0])
    inner_bdry = np.asarray(inner_bdry, dtype=np.float64)
    el_used = np.asarray(fn_elements_in_region({"nds": nds, "els": els}, inner_bdry), dtype=bool)
    main_int_el_i = np.nonzero(el_used)[0] + 1
    start = None
    for i in range(1, 5):
        el_i, bdry_nds = _adjacent_els(els, el_used, ~el_used)
        bdry_lyrs[bdry_nds - 1] = i
        el_used = el_used | el_i
        if i == 3:  # the third layer starts the absorbing region
            start = (nds[bdry_nds - 1] if abs_layer_start_bdry is None
                     else np.asarray(abs_layer_start_bdry, dtype=np.float64))
    # the main model's interface elements go; they are made again below
    keep = typ != "ASI2D2"
    els, mat, el_abs, typ, el_used = els[keep], mat[keep], el_abs[keep], typ[keep], el_used[keep]
    cand = ~el_used
    if cand.any():
        el_abs[cand] = fn_dist_point_to_bdry_2D(fn_calc_element_centres(nds, els[cand]),
                                                start) / abs_layer_thick
    keep = el_abs <= 1
    els, mat, el_abs, typ = els[keep], mat[keep], el_abs[keep], typ[keep]
    nds2, els2, old_nds, _ = fn_remove_unused_nodes(nds, els)
    dm = {"nds": nds2, "els": els2, "el_mat_i": mat, "el_abs_i": el_abs, "el_typ_i": typ,
          "bdry_lyrs": bdry_lyrs[old_nds - 1], "main_nd_i": old_nds,
          "main_int_el_i": main_int_el_i, "inner_bndry_pts": inner_bdry}
    dm = fn_add_fluid_solid_interface_els(dm, matls)
    free_ed = fn_find_free_edges(dm["els"]).ravel(order="F")
    dm["outer_bndry_pts"] = np.stack([dm["nds"][free_ed - 1, 0], dm["nds"][free_ed - 1, 1]], axis=1)
    dm["int_el_i"] = fn_elements_in_region(dm, inner_bdry)
    return dm


def _main_subdomain_mappings(dom_mod, main_res):
    """subdoms/fn_run_subdomain_model.m:119-145: the main model's monitored
    entries on the subdomain's boundary nodes (mn_res_i, 0-based, in the
    main monitor order), their global DOFs in the main model (gl_i, 1-based),
    and the subdomain node, DOF and layer of each."""
    lyrs = np.asarray(dom_mod["bdry_lyrs"])
    bdry_nds = np.nonzero(lyrs > 0)[0] + 1
    nb_m = np.asarray(dom_mod["main_nd_i"])[bdry_nds - 1]
    mon_nds = np.asarray(main_res["mon_nds"])
    mn_res_i = np.nonzero(np.isin(mon_nds, nb_m))[0]
    gl_i = np.asarray(main_res["dsp_gi"])[mn_res_i]
    Nb_m = mon_nds[mn_res_i]
    # interp1(nb_m, bdry_nds, Nb_m, 'nearest'): every Nb_m is one of nb_m
    sub_of = dict(zip(nb_m.tolist(), bdry_nds.tolist()))
    b_nds = np.asarray([sub_of[int(x)] for x in Nb_m], dtype=np.int64)
    return mn_res_i, gl_i, b_nds, np.asarray(main_res["mon_dfs"])[mn_res_i], lyrs[b_nds - 1]


def fn_run_subdomain_model(main, fe_options=None):
    """subdoms/fn_run_subdomain_model.m:1-116 after fn_run_main_model. For
    each subdomain: the main model's boundary impulse responses of every
    transmitter are convolved with the input signal and turned into the
    'in' forces on boundary layers 2 and 3 (fn_convert_disps_to_forces_v2 on
    the main model's matrices of the boundary DOFs); the subdomain model is
    marched for all transmitters at once (bfe_run_batched: one source column
    per transmitter, one signal row per force DOF, 64 columns per pass)
    where the reference runs one FE step per transmitter; the marched
    boundary displacements give the 'out' forces, which, convolved with each
    receiver's impulse response at those DOFs and summed, are added to the
    pristine FMC: main['doms'][d]['res']['fmc']. With finite
    field_output_every_n_frames the transmitters run one by one through
    fn_FE_entry_point, keeping each one's field frames."""
    from .driver import fn_FE_entry_point, fn_nds_and_dfs_to_gi
    n_tr = len(main["trans"])
    o = set_default_fields(dict(fe_options or {}), {
        "doms_to_run": [], "tx_trans": [], "rx_trans": [],
        "field_output_every_n_frames": np.inf, "use_gpu_if_available": 1, "dof_to_use": [],
        "solver_mode": "vel at curent time step"})
    typs = sorted(set(np.asarray(main["mod"]["el_typ_i"], dtype=object).tolist()))
    o["dof_to_use"] = find_dof_in_use_and_max_dof_per_el(typs, o["dof_to_use"])[0].tolist()
    tx = list(o["tx_trans"]) or list(range(1, n_tr + 1))
    rx = list(o["rx_trans"]) or list(range(1, n_tr + 1))
    doms = list(o["doms_to_run"]) or list(range(1, len(main["doms"]) + 1))
    time = np.asarray(main["inp"]["time"], dtype=np.float64)
    sig = np.asarray(main["inp"]["sig"], dtype=np.float64).ravel()
    n_t = time.size
    dt = time[1] - time[0]
    res_m = main["res"]
    K, C, M = res_m["mats"]["K"], res_m["mats"]["C"], res_m["mats"]["M"]
    mn_all = np.stack([np.asarray(res_m["mon_nds"]), np.asarray(res_m["mon_dfs"])], axis=1)
    field = np.isfinite(o["field_output_every_n_frames"])
    main = dict(main)
    main["doms"] = [dict(d) for d in main["doms"]]
    for d in doms:
        dom = main["doms"][d - 1]
        dm = dom["mod"]
        mn_res_i, gl_i, b_nds, b_dfs, b_lyrs = _main_subdomain_mappings(dm, res_m)
        g0 = gl_i - 1
        K_sub, C_sub, M_sub = K[g0][:, g0], C[g0][:, g0], M[g0][:, g0]
        # incident boundary displacements of every transmitter, convolved with
        # the input signal in one batched convolution (fn_run_subdomain_model.m:46)
        inc = np.concatenate([np.asarray(res_m["trans"][t - 1]["dsps"])[mn_res_i] for t in tx])
        inc = fn_convolve(inc, sig, 2).reshape(len(tx), mn_res_i.size, n_t)
        steps, frce_set = [], None
        for si, t in enumerate(tx):
            frcs, frce_set, _, _ = fn_convert_disps_to_forces_v2(K_sub, C_sub, M_sub, dt, inc[si],
                                                                 b_lyrs, "in", o["solver_mode"])
            steps.append({"load": {"frc_nds": b_nds[frce_set], "frc_dfs": b_dfs[frce_set],
                                   "frcs": frcs, "time": time},
                          "mon": {"nds": b_nds, "dfs": b_dfs,
                                  "field_output_every_n_frames": o["field_output_every_n_frames"]}})
        if field:
            res = fn_FE_entry_point(dm, main["matls"], steps, o)
            dsps = [r["dsps"] for r in res]
        else:
            dsps = _march_steps_batched(dm, main["matls"], steps, o, time)
        out = {"fmc": {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                       for k, v in res_m["fmc"].items()}}
        if field:
            out["trans"] = [dict() for _ in range(n_tr)]
            for si, t in enumerate(tx):
                out["trans"][t - 1] = {"fld": res[si]["fld"], "fld_time": res[si]["fld_time"]}
        fmc = out["fmc"]
        mn_nds = np.asarray(dm["main_nd_i"])[b_nds[frce_set] - 1]
        sel = _rows_in(mn_all, np.stack([mn_nds, b_dfs[frce_set]], axis=1))
        for si, t in enumerate(tx):
            frcs, _, _, _ = fn_convert_disps_to_forces_v2(K_sub, C_sub, M_sub, dt, dsps[si], b_lyrs,
                                                          "out", o["solver_mode"])
            # every receiver's impulse responses at the force DOFs, convolved
            # row by row with the forces and summed (:92-102)
            R = np.concatenate([np.asarray(res_m["trans"][r - 1]["dsps"])[sel] for r in rx])
            G = np.tile(frcs, (len(rx), 1))
            scat = fn_convolve(R, G, 2).reshape(len(rx), frcs.shape[0], n_t).sum(axis=1)
            for ri, r in enumerate(rx):
                k = np.nonzero((np.asarray(fmc["tx"]) == t) & (np.asarray(fmc["rx"]) == r))[0]
                fmc["time_data"][:, k] += scat[ri][:, None]
        dom["res"] = out
    return main


def _rows_in(a, b):
    """ismember(a, b, 'rows') for integer pairs."""
    kb = {tuple(x) for x in np.asarray(b).tolist()}
    return np.asarray([tuple(x) in kb for x in np.asarray(a).tolist()], dtype=bool)


def _march_steps_batched(dom_mod, matls, steps, o, time):
    """The load steps of one subdomain, all transmitters marched together:
    the model assembled once, one source column per step, one signal row per
    force DOF; returns each step's monitored displacements (rows in the
    steps' monitor order, as fn_BristolFE_v2 returns res.dsps)."""
    from .driver import fn_nds_and_dfs_to_gi
    opts = set_default_fields(dict(o), {"damping_power_law": 3, "max_damping": 3.1415e7,
                                        "max_stiffness_reduction": 0.01,
                                        "solver_precision": "double"})
    dm = dict(dom_mod)
    if "el_abs_i" not in dm:
        dm["el_abs_i"] = np.zeros(len(dm["el_mat_i"]))
    K, C, M, gl = fn_build_global_matrices_v5(dm["nds"], dm["els"], dm["el_mat_i"], dm["el_abs_i"],
                                              dm["el_typ_i"], matls, opts)
    n_t = time.size
    mon_gi, _, _, valid_m = fn_nds_and_dfs_to_gi(steps[0]["mon"]["nds"], steps[0]["mon"]["dfs"], gl)
    fd, fs, rows = [], [], []
    for s, st in enumerate(steps):
        gi, _, _, valid = fn_nds_and_dfs_to_gi(st["load"]["frc_nds"], st["load"]["frc_dfs"], gl)
        fr = np.atleast_2d(np.asarray(st["load"]["frcs"], dtype=np.float64))
        fd.append(gi - 1)
        fs.append(np.full(gi.size, s))
        rows.append(fr[valid] if fr.shape[0] > 1 else np.repeat(fr, gi.size, axis=0))
    fd, fs, F = np.concatenate(fd), np.concatenate(fs), np.concatenate(rows)
    plan = make_plan(K, C, M, time[1] - time[0], opts["solver_mode"], opts["solver_precision"])
    out = []
    try:
        nrow = int(np.nonzero(valid_m)[0].max()) + 1 if valid_m.any() else 0
        S = len(steps)
        for c0 in range(0, S, 64):
            c1 = min(c0 + 64, S)
            sel = (fs >= c0) & (fs < c1)
            h, _ = plan.run_batched(n_t, c1 - c0, fd[sel], fs[sel] - c0, F[sel], mon_gi - 1)
            for s in range(c1 - c0):
                d = np.zeros((nrow, n_t))
                d[np.nonzero(valid_m)[0]] = h[:, s, :]
                out.append(d)
    finally:
        plan.close()
    return out

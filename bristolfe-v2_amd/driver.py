"""Solver dispatch and FE driver: the reference's user-facing API.

  fn_FE_entry_point(mod, matls, steps, fe_options)   code/fn_FE_entry_point.m:1-64
  fn_BristolFE_v2(mod, matls, steps, fe_options)     code/fn_BristolFE_v2.m:1-168

``mod``, ``matls`` and ``steps`` keep the reference's field names, as dicts:
  mod   = {"nds", "els", "el_mat_i", ["el_abs_i"], ["el_typ_i"]}   (1-based)
  matls = [{"rho", "D", "el_typ", ...}, ...]    (matls[i-1] is material i)
  steps = [{"load": {"time", "frcs", "frc_nds", "frc_dfs", ["wts"]},
            "mon": {"nds", "dfs"}}, ...]      (a single dict is accepted)
Results are a list of dicts with the reference's fields
(dsps, dsp_gi, valid_mon_dsps, frcs, frc_gi, fld, fld_time).

``return_mats=True`` stands for MATLAB's second output ``[res, mats]``;
``steps=None`` returns the merged options (the options query,
fn_BristolFE_v2.m:33-37).
"""
from __future__ import annotations

import time as _time

import numpy as np

from . import _lib
from .assembly import fn_build_global_matrices_v5
from .field_output import fn_get_field_output
from .options import BRISTOLFE_DEFAULTS, ENTRY_POINT_DEFAULTS, set_default_fields
from .ebe import element_operator
from .solver import field_frames, fn_explicit_dynamic_solver_v6, make_plan


def fn_nds_and_dfs_to_gi(nds, dfs, gl_lookup):
    """(node, dof) -> 1-based global index (fn_BristolFE_v2.m:159-168). Pairs
    whose gl_lookup entry is 0 (a DOF the node does not carry) are dropped,
    as there; a node or DOF number outside gl_lookup raises IndexError, as
    MATLAB's gl_lookup(nds(i), dfs(i)) does."""
    nds = np.asarray(nds, dtype=np.int64).ravel()
    dfs = np.asarray(dfs, dtype=np.int64).ravel()
    if nds.size != dfs.size:
        raise ValueError("nds and dfs must have the same number of entries")
    bad = (nds < 1) | (nds > gl_lookup.shape[0]) | (dfs < 1) | (dfs > gl_lookup.shape[1])
    if bad.any():
        i = int(np.nonzero(bad)[0][0])
        raise IndexError(f"(node {nds[i]}, dof {dfs[i]}) is outside gl_lookup "
                         f"({gl_lookup.shape[0]} nodes x {gl_lookup.shape[1]} DOFs)")
    gi = gl_lookup[nds - 1, dfs - 1].astype(np.int64)
    valid = gi > 0
    return gi[valid], nds[valid], dfs[valid], valid


def device_field_kind(field_output_type, dim):
    """(type, is_vel, comp) of the device reduction for a field_output_type
    string (fn_get_field_output.m:28-79; frames as v6:64-69), None for
    'raw(u)' / 'raw(v)' (the nodal field itself) and for 'curl' on a 3D model
    (the host path raises the reference's error)."""
    t = field_output_type
    if t in ("raw(u)", "raw(v)"):
        return None
    if t in ("mean(u1)", "mean(u2)", "mean(u3)", "mean(v1)", "mean(v2)", "mean(v3)"):
        return (_lib.FIELD_MEAN, t[5] == "v", int(t[6]) - 1)
    if t in ("div(u)", "div(v)"):
        return (_lib.FIELD_DIV, t[4] == "v", 0)
    if t in ("curl(u)", "curl(v)"):
        return (_lib.FIELD_CURL, t[5] == "v", 0) if dim == 2 else None
    return (_lib.FIELD_KE, True, 0)  # the reference's 'otherwise' branch (:65)


def fn_FE_entry_point(mod, matls, steps, fe_options=None, return_mats=False, verbose=False):
    fe_options = set_default_fields(fe_options, ENTRY_POINT_DEFAULTS)
    solver = fe_options["solver"]
    if solver == "BristolFE":
        fn = fn_BristolFE_v2
    elif solver in ("pogo", "Abaqus"):
        raise NotImplementedError(f"external solver {solver!r} is not part of this library")
    else:
        raise ValueError("Invalid FE solver")
    t1 = _time.perf_counter()
    out = fn(mod, matls, steps, fe_options, return_mats=return_mats)
    if verbose:
        print(f"FE solver ({solver}) completed in {_time.perf_counter() - t1:.2f} secs")
    return out


def fn_BristolFE_v2(mod, matls, steps, fe_options=None, return_mats=False):
    fe_options = set_default_fields(fe_options, BRISTOLFE_DEFAULTS)
    if steps is None or (isinstance(steps, (list, tuple)) and len(steps) == 0):
        return fe_options  # options query

    mod = dict(mod)
    if "el_typ_i" not in mod:
        mod["el_typ_i"] = np.asarray([matls[i - 1]["el_typ"] for i in np.asarray(mod["el_mat_i"])],
                                     dtype=object)
    if "el_abs_i" not in mod:
        mod["el_abs_i"] = np.zeros(np.asarray(mod["el_mat_i"]).shape)

    K, C, M, gl_lookup = fn_build_global_matrices_v5(
        mod["nds"], mod["els"], mod["el_mat_i"], mod["el_abs_i"], mod["el_typ_i"], matls,
        fe_options)
    mats = {"K": K, "C": C, "M": M, "gl_lookup": gl_lookup}

    if isinstance(steps, dict):
        steps = [steps]
    if fe_options.get("dynamic_solver_version", "v6") not in ("v6", "hip"):
        raise NotImplementedError("only the v6 march (on the GPU) is provided")

    res = []
    plans = {}
    elem_op = None  # element-by-element form of K, built on first use
    try:
        for s, step in enumerate(steps):
            r = {}
            load = step.get("load", {})
            mon = step.get("mon", {})
            if "time" not in load:
                res.append(r)
                continue
            t = np.asarray(load["time"], dtype=np.float64).ravel()
            if "frcs" in load:
                frc_gi, _, _, valid_f = fn_nds_and_dfs_to_gi(load["frc_nds"], load["frc_dfs"], gl_lookup)
                frcs = np.asarray(load["frcs"], dtype=np.float64)
                if frcs.ndim == 1:
                    frcs = frcs[None, :]
                if "wts" in steps[0]["load"]:  # weights come from steps{1} (:79-80)
                    frcs = np.asarray(steps[0]["load"]["wts"], dtype=np.float64).reshape(-1, 1) @ frcs
                if frcs.shape[0] > 1:
                    frcs = frcs[valid_f]
            else:
                frc_gi, frcs = np.zeros(0, np.int64), np.zeros((0, t.size))
            if "dsps" in load:
                raise NotImplementedError(
                    "prescribed displacements: the v6 solver cannot run them (u_previous undefined)")
            if "nds" in mon:
                hist_gi, _, _, valid_m = fn_nds_and_dfs_to_gi(mon["nds"], mon["dfs"], gl_lookup)
                r["valid_mon_dsps"] = valid_m
            else:
                hist_gi = np.zeros(0, np.int64)
            dt = t[1] - t[0]
            key = (float(dt), fe_options["solver_mode"], fe_options["solver_precision"])
            if key not in plans:
                operator = fe_options.get("stiffness_operator", "auto")
                if operator != "csr" and elem_op is None:
                    elem_op = element_operator(mod, matls, fe_options, gl_lookup, K.shape[0])
                plans[key] = make_plan(K, C, M, dt, fe_options["solver_mode"],
                                       fe_options["solver_precision"], element_op=elem_op,
                                       operator=operator)
            frames = field_frames(t.size, fe_options["field_output_every_n_frames"])
            el_fld = None
            ftype = fe_options["field_output_type"]
            kind = device_field_kind(ftype, np.asarray(mod["nds"]).shape[1])
            if frames.size and kind is not None and fe_options.get("device_field_output", True):
                # the per-element field reduced on the device (fn_get_field_output.m:28-79)
                plan = plans[key]
                if getattr(plan, "n_field_el", None) is None:
                    plan.set_field_elements(mod["els"], gl_lookup)
                if kind[0] in (_lib.FIELD_DIV, _lib.FIELD_CURL) and \
                        getattr(plan, "field_dim", None) is None:
                    plan.set_field_geometry(mod["nds"])
                F = np.asarray(frcs, np.float64)
                F = F[None, :] if F.ndim == 1 else F
                mon_dsps, el_fld, ti_start = plan.run_field(
                    t.size, np.asarray(frc_gi, np.int64) - 1, F if np.size(frc_gi) else np.zeros((1, t.size)),
                    np.asarray(hist_gi, np.int64) - 1, int(fe_options["field_output_every_n_frames"]),
                    *kind)
                fld = None
                fld_time = np.where(frames >= ti_start, t[frames - 1], 0.0)
            else:
                mon_dsps, fld, _, fld_time = fn_explicit_dynamic_solver_v6(
                    K, C, M, t, frc_gi, frcs, None, None, hist_gi,
                    fe_options["field_output_every_n_frames"], fe_options["use_gpu_if_available"],
                    fe_options["field_output_type"], fe_options["solver_mode"],
                    fe_options["solver_precision"], plan=plans[key])
            r["fld_time"] = fld_time
            if "nds" in mon:
                valid_m = r["valid_mon_dsps"]
                # res.dsps(valid, :) = mon_dsps: rows up to the last valid one (:127)
                nrows = int(np.nonzero(valid_m)[0].max()) + 1 if valid_m.any() else 0
                d = np.zeros((nrows, t.size))
                d[np.nonzero(valid_m)[0]] = mon_dsps
                r["dsps"] = d
                r["dsp_gi"] = hist_gi
            else:
                r["dsps"] = np.zeros((0, 0))
                r["dsp_gi"] = np.zeros(0, np.int64)
            r["frcs"] = np.zeros((0, 0))
            r["frc_gi"] = np.zeros(0, np.int64)
            if el_fld is not None:
                r["fld"] = el_fld
            else:
                r["fld"] = (fn_get_field_output(fld, mod, mats, fe_options["field_output_type"])
                            if fld.size else np.zeros((0, 0)))
            res.append(r)
    finally:
        for p in plans.values():
            p.close()
    if return_mats:
        return res, mats
    return res

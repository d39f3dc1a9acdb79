"""Per-element field output from nodal field frames.

Mirror of code/internal/fn_get_field_output.m:1-134:
  'KE' (default)       sqrt of the element mean of sum_dof |v^2 m| / 2  (:64-79)
  'mean(ui)','mean(vi)' element mean of component i                  (:35-37)
  'div(.)','curl(.)'    sum over element nodes of the component along the
                        unit vector from the element centre         (:38-58, :85-105)
  'raw(u)','raw(v)'     the nodal field itself (the reference returns the type
                        string here, :61-62; returning the data is deliberate)
"""
from __future__ import annotations

import numpy as np

from .mesh import fn_calc_element_centres


def _global_to_nodal(gv, gl_lookup):
    nn, nd = gl_lookup.shape
    out = np.zeros((nn, gv.shape[1], nd))
    for i in range(nd):
        j = gl_lookup[:, i]
        valid = j > 0
        out[valid, :, i] = gv[j[valid] - 1, :]
    return out


def _average_over_element_nodes(nv, els):
    sz = nv.shape
    flat = nv.reshape(sz[0], -1)
    ev = np.zeros((els.shape[0], flat.shape[1]))
    cnt = np.zeros(els.shape[0])
    for i in range(els.shape[1]):
        j = els[:, i]
        valid = j > 0
        ev[valid] += flat[j[valid] - 1]
        cnt += valid
    ev /= cnt[:, None]
    return ev.reshape((els.shape[0],) + sz[1:])


def _unit_vecs_div(mod):
    nds = np.asarray(mod["nds"], dtype=np.float64)
    els = np.asarray(mod["els"])
    valid = np.all(els > 0, axis=1)
    uv = nds[np.maximum(els, 1) - 1] - fn_calc_element_centres(nds, els)[:, None, :]
    uv = uv / np.sqrt((uv ** 2).sum(axis=2, keepdims=True))
    uv[~valid] = 0
    return uv


def fn_get_field_output(f_out, mod, mats, field_output_type="sqrt(KE)"):
    f_out = np.asarray(f_out)
    els = np.asarray(mod["els"])
    gl = mats["gl_lookup"]
    t = field_output_type
    if t in ("mean(u1)", "mean(u2)", "mean(u3)", "mean(v1)", "mean(v2)", "mean(v3)"):
        nv = _global_to_nodal(f_out, gl)
        return _average_over_element_nodes(nv[:, :, int(t[6]) - 1], els)
    if t in ("div(u)", "div(v)", "curl(u)", "curl(v)"):
        nv = _global_to_nodal(f_out, gl)
        uv = _unit_vecs_div(mod)
        if t.startswith("curl"):
            if np.asarray(mod["nds"]).shape[1] > 2:
                raise ValueError("Cannot output curl for 3D models")
            uv = np.stack([uv[:, :, 1], -uv[:, :, 0]], axis=2)
        v = np.zeros((els.shape[0], f_out.shape[1]))
        nc = uv.shape[2]
        for i in range(els.shape[1]):
            j = np.maximum(els[:, i], 1) - 1
            v += (nv[j, :, :nc] * uv[:, i, None, :]).sum(axis=2)
        return v
    if t in ("raw(u)", "raw(v)"):
        return f_out
    diag_m = mats["M"].diagonal()
    KE_g = np.abs(f_out ** 2 * diag_m[:, None]) / 2
    KE_n = _global_to_nodal(KE_g, gl).sum(axis=2)
    return np.sqrt(_average_over_element_nodes(KE_n, els))

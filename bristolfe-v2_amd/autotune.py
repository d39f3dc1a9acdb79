"""Autotuner of the multi-source march: picks the launch shape by timing.

Tunables (none changes results):
  rows    wavefronts per workgroup of the one-wavefront-per-row multi-source
          kernel: 1, 2, 4, 8
  rpw     matrix rows each wavefront handles together (interleaved): 1, 2, 4,
          8, 16
  remap   XCD-aware workgroup order on/off
  group   rows per gather group (one u1 gather serves the group's rows over
          the union of their columns): automatic, 1, 2, 4, 8
  tile    source columns marched together per pass: 64 or 32 (fewer columns
          per pass re-read the operator more often but keep more lanes busy
          per byte of u)
  operator  the CSR operator, or (with an element operator attached whose
          groups all have a mesh-pattern kernel) the multi-source element
          product with npw nodes per wavefront: 2, 4, 8, or (triangle
          patterns) the same kernel on the node-block form of the element
          operator
Each candidate marches a short probe (a prefix of the real time axis) on the
real operator; the fastest by HIP-event loop time wins. The search is staged:
rows x rpw first (XCD order on, automatic group, widest tile), then each
other group with one or two groups per wavefront over all rows, then the XCD
order, then the tile width, each at the best shape found so far, then the
element product's and the node-block form's nodes per wavefront.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from . import _lib


def run_tiled(plan, n_t, n_src, frc_dof, frc_src, frcs, hist_dof, tile):
    """Marches n_src columns in passes of `tile`; returns hist (q, n_src, n_t)
    and the summed loop time (ms). frcs holds one signal row for every
    forcing entry, or one row broadcast to all of them; each pass takes the
    rows of its own entries."""
    q = np.asarray(hist_dof).size
    frc_dof = np.asarray(frc_dof)
    frc_src = np.asarray(frc_src)
    frcs = np.atleast_2d(frcs)
    if frcs.shape[0] not in (1, frc_dof.size):
        raise ValueError("forcing functions must have 1 or p rows")
    if tile >= n_src:  # one pass: the library's output is the result
        h, _ = plan.run_batched(n_t, n_src, frc_dof, frc_src, frcs, hist_dof)
        return h, plan.stats()["loop_ms"]
    out = np.zeros((q, n_src, n_t))
    ms = 0.0
    for c0 in range(0, n_src, tile):
        c1 = min(c0 + tile, n_src)
        sel = (frc_src >= c0) & (frc_src < c1)
        rows = frcs if frcs.shape[0] == 1 else frcs[sel]
        h, _ = plan.run_batched(n_t, c1 - c0, frc_dof[sel], frc_src[sel] - c0, rows, hist_dof)
        out[:, c0:c1, :] = h
        ms += plan.stats()["loop_ms"]
    return out, ms


def tune_batched(plan, n_src, frc_dof, frc_src, frcs, hist_dof, probe_steps=120,
                 rows=(1, 2, 4, 8), rpws=(1, 2, 4, 8, 16), remaps=(1, 0), tiles=(64, 32),
                 groups=(1, 2, 4, 8), repeats=2, ebe_npw=(2, 4, 8)):
    """Returns (best, table): best = dict(rows, rpw, remap, group, tile,
    operator, npw, ms_per_step); group is the one the library used, operator
    "csr", "ebe" or "node" (npw its nodes per wavefront)."""
    frcs = np.atleast_2d(frcs)
    # the probe must include forced steps: start it at the first non-zero sample
    nz = np.nonzero(np.abs(frcs).sum(axis=0))[0]
    t0 = int(nz[0]) if nz.size else 0
    probe = frcs[:, :t0 + probe_steps]
    table = []

    def measure(r, w, x, g, tile, npw=0, op=_lib.OPERATOR_EBE):
        plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_BLOCK, r)
        plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_WAVE, w)
        plan.set_tuning(_lib.TUNE_XCD_REMAP, x)
        plan.set_tuning(_lib.TUNE_MULTI_GROUP, g)
        if npw:
            plan.set_tuning(_lib.TUNE_MULTI_NODES_PER_WAVE, npw)
        plan.set_tuning(_lib.TUNE_MULTI_OPERATOR, op if npw else _lib.OPERATOR_CSR)
        best = np.inf
        for _ in range(repeats):
            _, ms = run_tiled(plan, probe.shape[1], n_src, frc_dof, frc_src, probe, hist_dof, tile)
            best = min(best, ms)
        steps = max(int(plan.stats()["steps"]), 1)
        info = plan.info()
        used = info.get("multi_operator", 0) if npw else _lib.OPERATOR_CSR
        name = {_lib.OPERATOR_EBE: "ebe", _lib.OPERATOR_NODE: "node"}.get(used, "csr")
        rec = dict(rows=r, rpw=w, remap=x, group=int(info["multi_group"]), tile=tile,
                   operator=name, npw=int(npw) if name != "csr" else 0,
                   ms_per_step=best / steps)
        table.append(rec)
        return rec

    def keep(rec, best):
        return rec if rec["ms_per_step"] < best["ms_per_step"] else best

    widest = min(tiles[0], n_src)
    best = min((measure(r, w, remaps[0], 0, widest) for r, w in itertools.product(rows, rpws)),
               key=lambda d: d["ms_per_step"])
    for g in groups:  # each group at one and two groups per wavefront
        if g == best["group"]:
            continue
        for w in sorted({g, 2 * g} & set(rpws)):
            for r in rows:
                best = keep(measure(r, w, remaps[0], g, widest), best)
    for x in remaps[1:]:
        best = keep(measure(best["rows"], best["rpw"], x, best["group"], widest), best)
    for tile in tiles[1:]:
        if tile >= n_src:
            continue
        best = keep(measure(best["rows"], best["rpw"], best["remap"], best["group"], tile), best)
    info = plan.info()
    if info.get("operator_kind", 0) == 1 and info.get("ebe_pattern", 0) != 0:
        # the element product and its node-block form, where the plan can run them
        for op, name in ((_lib.OPERATOR_EBE, "ebe"), (_lib.OPERATOR_NODE, "node")):
            for npw in ebe_npw:
                rec = measure(best["rows"], best["rpw"], best["remap"], best["group"],
                              best["tile"], npw, op)
                if rec["operator"] != name:
                    break
                best = keep(rec, best)
    apply(plan, best)
    return best, table


def apply(plan, shape):
    """Sets a launch shape found by tune_batched on a plan."""
    plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_BLOCK, shape["rows"])
    plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_WAVE, shape["rpw"])
    plan.set_tuning(_lib.TUNE_XCD_REMAP, shape["remap"])
    plan.set_tuning(_lib.TUNE_MULTI_GROUP, shape["group"])
    op = {"ebe": _lib.OPERATOR_EBE, "node": _lib.OPERATOR_NODE}.get(shape.get("operator", "csr"),
                                                                    _lib.OPERATOR_CSR)
    plan.set_tuning(_lib.TUNE_MULTI_OPERATOR, op)
    if op != _lib.OPERATOR_CSR:
        plan.set_tuning(_lib.TUNE_MULTI_NODES_PER_WAVE, shape["npw"])

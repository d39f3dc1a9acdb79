"""Autotuner of the multi-source march: picks the launch shape by timing.

Tunables (none changes results):
  rows    wavefronts (= matrix rows) per workgroup of the one-wave-per-row
          multi-source kernel: 1, 2, 4, 8
  remap   XCD-aware workgroup order on/off
  tile    source columns marched together per pass: 16, 32 or 64 (fewer
          columns per pass re-read the operator more often but keep more
          lanes busy per byte of u)
Each candidate marches a short probe (a prefix of the real time axis) on the
real operator; the fastest by HIP-event loop time wins.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from . import _lib


def run_tiled(plan, n_t, n_src, frc_dof, frc_src, frcs, hist_dof, tile):
    """Marches n_src columns in passes of `tile`; returns hist (q, n_src, n_t)
    and the summed loop time (ms)."""
    q = np.asarray(hist_dof).size
    out = np.zeros((q, n_src, n_t))
    ms = 0.0
    frc_dof = np.asarray(frc_dof)
    frc_src = np.asarray(frc_src)
    for c0 in range(0, n_src, tile):
        c1 = min(c0 + tile, n_src)
        sel = (frc_src >= c0) & (frc_src < c1)
        h, _ = plan.run_batched(n_t, c1 - c0, frc_dof[sel], frc_src[sel] - c0, frcs, hist_dof)
        out[:, c0:c1, :] = h
        ms += plan.stats()["loop_ms"]
    return out, ms


def tune_batched(plan, n_src, frc_dof, frc_src, frcs, hist_dof, probe_steps=120,
                 rows=(1, 2, 4, 8), remaps=(1, 0), tiles=(64, 32, 16), repeats=2):
    """Returns (best, table): best = dict(rows, remap, tile, ms_per_step)."""
    frcs = np.atleast_2d(frcs)
    # the probe must include forced steps: start it at the first non-zero sample
    nz = np.nonzero(np.abs(frcs).sum(axis=0))[0]
    t0 = int(nz[0]) if nz.size else 0
    probe = frcs[:, :t0 + probe_steps]
    table = []
    for r, x, tile in itertools.product(rows, remaps, tiles):
        if tile > n_src and tile != tiles[0]:
            continue
        plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_BLOCK, r)
        plan.set_tuning(_lib.TUNE_XCD_REMAP, x)
        best = np.inf
        for _ in range(repeats):
            _, ms = run_tiled(plan, probe.shape[1], n_src, frc_dof, frc_src, probe, hist_dof,
                              min(tile, n_src))
            best = min(best, ms)
        table.append(dict(rows=r, remap=x, tile=min(tile, n_src), ms_per_step=best / probe_steps))
    best = min(table, key=lambda d: d["ms_per_step"])
    plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_BLOCK, best["rows"])
    plan.set_tuning(_lib.TUNE_XCD_REMAP, best["remap"])
    return best, table

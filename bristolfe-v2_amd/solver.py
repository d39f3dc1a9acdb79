"""Explicit dynamic solver with the v6 signature, marching on the GPU.

Mirror of code/internal/fn_explicit_dynamic_solver_v6.m:1-221:

  hist, fld, frc_out, fld_time = fn_explicit_dynamic_solver_v6(
      K, C, M, time, forcing_indices, forcing_functions, disp_indices,
      disp_functions, history_indices, field_output_every_n_frames,
      use_gpu_if_present=1, field_output_type='KE',
      solver_mode='vel at curent time step', solver_precision='double')

Indices are 1-based global DOF numbers (MATLAB convention); K, C, M are
scipy sparse matrices. The march runs in ``libbfe_hip.so`` (C ABI,
include/bfe.h); there is no CPU path, so a missing library or device raises.

Differences from v6, all deliberate:
  * prescribed displacements (v6:102-110, 189-192) raise NotImplementedError:
    v6 references the undefined ``u_previous`` there and cannot run them;
  * the correctly spelled 'vel at current time step' is accepted as an alias
    of the reference's 'vel at curent time step' (v6:47 vs :135);
  * ``use_gpu_if_present`` = 0 is ignored: the march always runs on the GPU.
"""
from __future__ import annotations

import math
import warnings

import numpy as np

from . import _lib
from .options import SOLVER_MODES

VEL_FIELD_TYPES_U = {"mean(u1)", "mean(u2)", "mean(u3)", "curl(u)", "div(u)", "raw(u)"}


def _mode(solver_mode: str) -> int:
    try:
        return SOLVER_MODES[solver_mode]
    except KeyError:
        raise ValueError(f"unknown solver_mode {solver_mode!r}") from None


def _precision(p: str) -> int:
    p = str(p).lower()
    if p == "double":
        return _lib.PREC_DOUBLE
    if p == "single":
        return _lib.PREC_SINGLE
    raise ValueError(f"unknown solver_precision {p!r}")


def lumped_mass(M) -> np.ndarray:
    """m = sum(M) (column sums, v6:122)."""
    return np.asarray(M.sum(axis=0)).ravel()


# "auto" uses the element operator from this many DOF up: below it a step is
# bound by the latency of one kernel (about 4.5 us on the 40,602-DOF plate
# with either operator) and the CSR step is as fast.
AUTO_EBE_MIN_DOF = 1_000_000


def make_plan(K, C, M, dt, solver_mode="vel at curent time step", solver_precision="double",
              device=0, element_op=None, operator="auto") -> _lib.Plan:
    """Plan of the march. With an ElementOperator (bristolfe_v2_amd.ebe) the
    single-source march uses the element-by-element step: operator "ebe"
    requires it, "auto" uses it from AUTO_EBE_MIN_DOF DOF up and falls back to
    the CSR step when the library rejects it, "csr" ignores it."""
    if K.shape[0] != K.shape[1]:
        raise ValueError("K must be square matrix")
    if C.shape[0] != C.shape[1]:
        raise ValueError("C must be square matrix")
    if M.shape[0] != M.shape[1]:
        raise ValueError("M must be square matrix")
    if operator not in ("auto", "csr", "ebe"):
        raise ValueError(f"unknown stiffness operator {operator!r}")
    if operator == "ebe" and element_op is None:
        raise ValueError("operator 'ebe' needs an element operator for this model")
    plan = _lib.Plan.from_scipy(K, lumped_mass(M), C, dt, _mode(solver_mode),
                                _precision(solver_precision), device)
    if operator == "auto" and K.shape[0] < AUTO_EBE_MIN_DOF:
        element_op = None
    if element_op is not None and operator != "csr":
        try:
            plan.set_element_operator(element_op)
        except _lib.BfeError as e:
            if operator == "ebe" or e.status != _lib.BFE_ERR_ARG:
                plan.close()
                raise
    return plan


def field_frames(n_t: int, every) -> np.ndarray:
    """1-based frame steps 1:N:n_t (v6:113); empty for N = inf."""
    if every is None or (isinstance(every, float) and math.isinf(every)) or every <= 0:
        return np.zeros(0, dtype=np.int64)
    return np.arange(1, n_t + 1, int(every))


def fn_explicit_dynamic_solver_v6(K, C, M, time, forcing_indices, forcing_functions,
                                  disp_indices, disp_functions, history_indices,
                                  field_output_every_n_frames, use_gpu_if_present=1,
                                  field_output_type="KE",
                                  solver_mode="vel at curent time step",
                                  solver_precision="double", plan: _lib.Plan | None = None):
    time = np.asarray(time, dtype=np.float64).ravel()
    n_t = time.size
    dt = time[1] - time[0]
    if disp_indices is not None and np.size(disp_indices) > 0:
        raise NotImplementedError(
            "prescribed displacements: the v6 solver cannot run them (u_previous undefined)")
    own = plan is None
    if own:
        plan = make_plan(K, C, M, dt, solver_mode, solver_precision)
    try:
        frc = np.asarray(forcing_indices if forcing_indices is not None else [], dtype=np.int64).ravel() - 1
        F = np.asarray(forcing_functions if forcing_functions is not None else np.zeros((0, n_t)),
                       dtype=np.float64)
        if F.ndim == 1:
            F = F[None, :]
        hidx = np.asarray(history_indices if history_indices is not None else [], dtype=np.int64).ravel() - 1
        frames = field_frames(n_t, field_output_every_n_frames)
        is_vel = 0 if field_output_type in VEL_FIELD_TYPES_U else 1
        every = int(field_output_every_n_frames) if frames.size else 0
        hist, fld, ti_start = plan.run(n_t, frc, F if frc.size else np.zeros((1, n_t)), hidx,
                                       every, is_vel)
        st = plan.stats()
        if st["nonfinite"]:
            warnings.warn(f"explicit march produced {st['nonfinite']} non-finite history values "
                          "(time step above the stability limit?)", RuntimeWarning)
    finally:
        if own:
            plan.close()
    history_output = hist if hidx.size else np.zeros((0, n_t))
    if frames.size:
        fld_time = np.where(frames >= ti_start, time[frames - 1], 0.0)
    else:
        fld = np.zeros((K.shape[0], 0))
        fld_time = np.zeros(0)
    return history_output, fld, np.zeros((0, n_t)), fld_time

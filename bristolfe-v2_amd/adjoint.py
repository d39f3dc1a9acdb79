"""Differentiable march: gradients of an A-scan misfit with respect to a
per-element stiffness scale s_e and density rho_e.

No counterpart in the reference (which has no adjoint). The forward model is
the reference's: fn_build_global_matrices_v5.m builds K, C, M and
fn_explicit_dynamic_solver_v6.m marches them. Parameters enter as

    K(s)   = K0 + sum_e (s_e - 1) K_e
    M(rho) = M0 + sum_e (rho_e - rho0_e) Mbar_e
    C(rho) = C0 + sum_e mu_e (rho_e - rho0_e) Mbar_e

with K_e the element stiffness including its absorbing-layer reduction
(fn_build_global_matrices_v5.m:116), Mbar_e = M_e / rho0_e the lumped mass per
unit density and mu_e = a_e^p * max_damping its absorbing damping per unit
mass (:115). At s = 1, rho = rho0 the matrices are exactly the assembled
ones. Parameter elements are the solid ones (CPE3, C3D4).

The march and its adjoint run in libbfe_hip.so: ``bfe_run_checkpointed``
(one transmitter) or ``bfe_run_checkpointed_batched`` (up to 32 transmitters
in one multi-source march) keeps checkpoints every ``ckpt_every`` steps,
``bfe_adjoint`` marches the adjoint backward (fused with the recomputation
of the checkpointed segments) and accumulates the gradients in FP64.

``MarchFunction`` wraps this as a torch.autograd.Function on CPU float64
tensors (torch only does the bookkeeping; the GPU work is in the library).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib
from .assembly import fn_build_global_matrices_v5
from .elements import element_function
from .options import SOLVER_MODES, set_default_fields
from .solver import _precision

PARAM_TYPES = ("CPE3", "C3D4")
_NODES = {"CPE3": 3, "C3D4": 4}


class ParametrizedModel:
    """A model whose solid elements carry (s_e, rho_e)."""

    def __init__(self, mod, matls, fe_options=None):
        o = set_default_fields(dict(fe_options or {}), {
            "dof_to_use": [], "damping_power_law": 3, "max_damping": 3.1415e7,
            "max_stiffness_reduction": 0.01, "solver_mode": "vel at curent time step",
            "solver_precision": "double"})
        self.fe_options = o
        mod = dict(mod)
        if "el_typ_i" not in mod:
            mod["el_typ_i"] = np.asarray([matls[i - 1]["el_typ"] for i in mod["el_mat_i"]], dtype=object)
        if "el_abs_i" not in mod:
            mod["el_abs_i"] = np.zeros(len(mod["el_mat_i"]))
        self.mod, self.matls = mod, matls
        self.K0, self.C0, self.M0, self.gl = fn_build_global_matrices_v5(
            mod["nds"], mod["els"], mod["el_mat_i"], mod["el_abs_i"], mod["el_typ_i"], matls, o)
        self.n = self.K0.shape[0]
        typ = np.asarray(mod["el_typ_i"], dtype=object)
        mat = np.asarray(mod["el_mat_i"])
        a = np.asarray(mod["el_abs_i"], dtype=np.float64)
        pl, md = float(o["damping_power_law"]), float(o["max_damping"])
        lmsr = np.log(float(o["max_stiffness_reduction"]))
        kmax = 0
        groups = []
        for t in PARAM_TYPES:
            for mi in np.unique(mat[typ == t]):
                sel = np.nonzero((typ == t) & (mat == mi))[0]
                m = matls[int(mi) - 1]
                rho = float(m["density"] if "density" in m else m["rho"])
                e = np.asarray(mod["els"])[sel, :_NODES[t]] - 1
                K, _, M, ln, ld = element_function(t)(np.asarray(mod["nds"], dtype=np.float64), e,
                                                      np.asarray(m["D"], dtype=np.float64), rho,
                                                      o["dof_to_use"])
                K = K * np.exp(lmsr * a[sel] ** (pl + 1))[:, None, None]
                dofs = self.gl[e[:, np.asarray(ln) - 1], np.asarray(ld) - 1] - 1
                k = len(ln)
                groups.append((sel, dofs, K, np.diagonal(M, axis1=1, axis2=2) / rho,
                               a[sel] ** pl * md, np.full(sel.size, rho)))
                kmax = max(kmax, k)
        if not groups:
            raise ValueError("the model has no solid (CPE3/C3D4) elements to differentiate")
        n_el = sum(g[0].size for g in groups)
        self.el_index = np.concatenate([g[0] for g in groups])  # model element numbers (0-based)
        self.el_dofs = np.full((n_el, kmax), -1, dtype=np.int32)
        self.el_K = np.zeros((n_el, kmax, kmax))
        self.el_m = np.zeros((n_el, kmax))
        self.el_mu = np.zeros(n_el)
        self.rho0 = np.zeros(n_el)
        o0 = 0
        for sel, dofs, K, mb, mu, r0 in groups:
            k = dofs.shape[1]
            s = slice(o0, o0 + sel.size)
            self.el_dofs[s, :k] = dofs
            self.el_K[s, :k, :k] = K
            self.el_m[s, :k] = mb
            self.el_mu[s] = mu
            self.rho0[s] = r0
            o0 += sel.size
        self.n_el = n_el
        self._pattern()

    def _pattern(self):
        """Coordinate lists to scatter element data into global matrices."""
        k = self.el_dofs.shape[1]
        valid = self.el_dofs >= 0
        r = np.broadcast_to(self.el_dofs[:, :, None], (self.n_el, k, k))
        c = np.broadcast_to(self.el_dofs[:, None, :], (self.n_el, k, k))
        vv = valid[:, :, None] & valid[:, None, :]
        self._kr, self._kc, self._kmask = r[vv], c[vv], vv
        self._md, self._mmask = self.el_dofs[valid], valid

    def matrices(self, s=None, rho=None):
        """K(s), C(rho), M(rho) as scipy CSR (exactly the assembled ones at s = 1, rho = rho0)."""
        K, C, M = self.K0, self.C0, self.M0
        if s is not None:
            ds = np.asarray(s, dtype=np.float64) - 1.0
            if np.any(ds):
                v = (self.el_K * ds[:, None, None])[self._kmask]
                K = (K + sp.csr_matrix((v, (self._kr, self._kc)), shape=K.shape)).tocsr()
        if rho is not None:
            dr = np.asarray(rho, dtype=np.float64) - self.rho0
            if np.any(dr):
                dm = np.bincount(self._md, weights=(self.el_m * dr[:, None])[self._mmask],
                                 minlength=self.n)
                dc = np.bincount(self._md, weights=(self.el_m * (dr * self.el_mu)[:, None])[self._mmask],
                                 minlength=self.n)
                M = (M + sp.diags(dm)).tocsr()
                C = (C + sp.diags(dc)).tocsr()
        return K, C, M

    def plan(self, dt, s=None, rho=None, device=0):
        K, C, M = self.matrices(s, rho)
        o = self.fe_options
        p = _lib.Plan.from_scipy(K, np.asarray(M.sum(axis=0)).ravel(), C, dt,
                                 SOLVER_MODES[o["solver_mode"]], _precision(o["solver_precision"]),
                                 device)
        p.set_gradient_elements(self.el_dofs, self.el_K, self.el_m, self.el_mu)
        return p


class Shots:
    """Transmit/receive set: per transmitter the forcing DOFs (0-based) and a
    signal; the same receiver DOFs (0-based) for every transmitter."""

    def __init__(self, tx_dofs, signal, rx_dofs, n_t, dt, ckpt_every=250):
        self.tx = [np.atleast_1d(np.asarray(t, dtype=np.int64)) for t in tx_dofs]
        self.signal = np.atleast_2d(np.asarray(signal, dtype=np.float64))
        self.rx = np.asarray(rx_dofs, dtype=np.int64).ravel()
        self.n_t, self.dt, self.ckpt_every = int(n_t), float(dt), int(ckpt_every)


def forward(model: ParametrizedModel, shots: Shots, s=None, rho=None, keep=False):
    """A-scans (n_tx, n_rx, n_t). With keep=True also returns the plan holding
    every transmitter's checkpoints (slot = transmitter index) and the summed
    HIP-event loop time of the forward marches."""
    plan = model.plan(shots.dt, s, rho)
    out = np.zeros((len(shots.tx), shots.rx.size, shots.n_t))
    ms = 0.0
    try:
        for i, tx in enumerate(shots.tx):
            h, _ = plan.run_checkpointed(i, shots.n_t, tx, shots.signal, shots.rx, shots.ckpt_every)
            out[i] = h
            ms += plan.stats()["loop_ms"]
    except Exception:
        plan.close()
        raise
    if keep:
        return out, plan, ms
    plan.close()
    return out


def _entry_signals(shots: Shots, txs):
    """Signal rows of a multi-source march over the transmitters txs: the one
    shared row, or, when every transmitter forces one DOF per signal row, the
    rows repeated for each transmitter (the forcing entries' order)."""
    sig = shots.signal
    if sig.shape[0] == 1:
        return sig
    if any(t.size != sig.shape[0] for t in txs):
        raise ValueError("signal rows must be 1 or one per forcing DOF of every transmitter")
    return np.concatenate([sig] * len(txs))


def forward_batched(model: ParametrizedModel, shots: Shots, s=None, rho=None, keep=False):
    """A-scans (n_tx, n_rx, n_t) of all transmitters (at most 32) in one
    checkpointed multi-source march. With keep=True also returns the plan
    holding the checkpoints (slot 0) for backward_batched and the HIP-event
    loop time."""
    if len(shots.tx) > 32:
        raise ValueError("the batched adjoint takes at most 32 transmitters")
    plan = model.plan(shots.dt, s, rho)
    fi = np.concatenate(shots.tx)
    fs = np.concatenate([np.full(t.size, k) for k, t in enumerate(shots.tx)])
    sig = _entry_signals(shots, shots.tx)
    try:
        h, _ = plan.run_checkpointed_batched(0, shots.n_t, len(shots.tx), fi, fs, sig, shots.rx,
                                             shots.ckpt_every)
        ms = plan.stats()["loop_ms"]
    except Exception:
        plan.close()
        raise
    out = np.ascontiguousarray(np.transpose(h, (1, 0, 2)))
    if keep:
        return out, plan, ms
    plan.close()
    return out


def backward_batched(plan, grad_ascans):
    """dJ/ds, dJ/drho from dJ/dA-scans (n_tx, n_rx, n_t) of a forward_batched
    plan (consumes its checkpoints): one reverse sweep for all transmitters."""
    g = np.asfortranarray(np.transpose(np.asarray(grad_ascans, np.float64), (1, 0, 2)))
    gk, gm = plan.adjoint(0, g)
    return gk, gm, plan.stats()["loop_ms"]


def ascans_batched(model: ParametrizedModel, shots: Shots, s=None, rho=None):
    """A-scans (n_tx, n_rx, n_t) of all transmitters in one multi-source march
    (no checkpoints): the cheap evaluator for finite differences."""
    K, C, M = model.matrices(s, rho)
    o = model.fe_options
    out = np.zeros((len(shots.tx), shots.rx.size, shots.n_t))
    for c0 in range(0, len(shots.tx), 64):
        grp = shots.tx[c0:c0 + 64]
        fi = np.concatenate(grp)
        fs = np.concatenate([np.full(t.size, k) for k, t in enumerate(grp)])
        with _lib.Plan.from_scipy(K, np.asarray(M.sum(axis=0)).ravel(), C, shots.dt,
                                  SOLVER_MODES[o["solver_mode"]], _precision(o["solver_precision"])) as p:
            h, _ = p.run_batched(shots.n_t, len(grp), fi, fs, _entry_signals(shots, grp),
                                 shots.rx)
        out[c0:c0 + len(grp)] = np.transpose(h, (1, 0, 2))
    return out


def backward(plan, grad_ascans):
    """dJ/ds, dJ/drho from dJ/dA-scans; consumes the plan's checkpoints."""
    gk = gm = None
    ms = 0.0
    for i in range(grad_ascans.shape[0]):
        a, b = plan.adjoint(i, grad_ascans[i])
        ms += plan.stats()["loop_ms"]
        gk = a if gk is None else gk + a
        gm = b if gm is None else gm + b
    return gk, gm, ms


def _torch():
    import torch  # imported lazily: only the autograd wrapper needs it
    return torch


def march_function():
    torch = _torch()

    class MarchFunction(torch.autograd.Function):
        """A-scans = march(s, rho); backward = adjoint march (CPU float64 tensors)."""

        @staticmethod
        def forward(ctx, s, rho, model, shots):
            # all transmitters in one checkpointed multi-source march when
            # they fit one (at most 32), else one march per transmitter
            ctx.batched = len(shots.tx) <= 32
            fwd = forward_batched if ctx.batched else forward
            a, plan, _ = fwd(model, shots, s.detach().cpu().numpy(), rho.detach().cpu().numpy(),
                             keep=True)
            ctx.plan = plan
            return torch.from_numpy(a)

        @staticmethod
        def backward(ctx, grad):
            plan = ctx.plan
            bwd = backward_batched if ctx.batched else backward
            try:
                gk, gm, _ = bwd(plan, grad.detach().cpu().numpy().astype(np.float64))
            finally:
                plan.close()
                ctx.plan = None
            return torch.from_numpy(gk), torch.from_numpy(gm), None, None

    return MarchFunction


def ascan_misfit(model, shots, data, s, rho):
    """J = 1/2 sum (A(s, rho) - data)^2 with autograd through the adjoint."""
    torch = _torch()
    A = march_function().apply(s, rho, model, shots)
    return 0.5 * torch.sum((A - torch.as_tensor(data)) ** 2)

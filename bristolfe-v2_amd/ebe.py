"""Element-by-element form of the global stiffness for structured meshes.

K as assembled by fn_build_global_matrices_v5.m:105-128 is a sum of element
matrices, each scaled by its absorbing-layer reduction exp(ln(msr) a^(p+1))
(:116). On the structured meshes the reference's meshers produce
(code/fn_rectangular_structured_mesh.m, and the tet split of
code/fn_3d_cubic_structured_mesh.m's cells) every element is a translate of
one of a few shapes, so the element matrices are a few templates times a
per-element scale:

    K = sum_e s_e P_e' T[tau_e] P_e,   tau_e = (material, shape)

``element_operator`` finds the templates (shapes are compared after rounding
the node offsets to 2^-24 of the largest element extent) and returns the
arrays that ``Plan.set_element_operator`` hands to the library, which checks
one step against the CSR operator before using it. Returns None when the
model does not have that form: more than one element type, DOFs not numbered
node-major with the same DOFs on every node, or too many templates.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .elements import element_function
from .options import set_default_fields

_NPE = {"CPE3": 3, "AC2D3": 3, "C3D4": 4}
_CHUNK = 1 << 20
_MAX_SLOTS = 64


@dataclass
class ElementOperator:
    n_nodes: int
    dpn: int
    el_nodes: np.ndarray   # (n_el, npe) int32, 0-based
    el_tmpl: np.ndarray    # (n_el,) int32
    el_scale: np.ndarray   # (n_el,) float64
    tmpl: np.ndarray       # (n_tmpl, npe*dpn, npe*dpn) float64

    @property
    def n_el(self) -> int:
        return int(self.el_nodes.shape[0])


def _unique_type(el_typ_i):
    a = np.asarray(el_typ_i, dtype=object).ravel()
    if a.size == 0 or not bool(np.all(a == a[0])):
        return None
    return a[0]


def element_operator(mod, matls, fe_options, gl_lookup, n_dof):
    """ElementOperator for the model, or None if it does not qualify."""
    o = set_default_fields(dict(fe_options or {}), {
        "dof_to_use": [], "damping_power_law": 3, "max_stiffness_reduction": 0.01})
    if "el_typ_i" in mod:
        typ = _unique_type(mod["el_typ_i"])
    else:
        typ = _unique_type([matls[i - 1]["el_typ"] for i in np.unique(mod["el_mat_i"])])
    if typ not in _NPE:
        return None
    npe = _NPE[typ]
    fn = element_function(typ)
    ln, ld = fn(None, None, None, None, o["dof_to_use"])
    ln, ld = np.asarray(ln), np.asarray(ld)
    if ln.size == 0 or ln.size % npe:
        return None
    dpn = ln.size // npe
    dofs = ld[:dpn]
    # local DOF order must be node-major with the same DOFs on every vertex
    if not (np.array_equal(ln, np.repeat(np.arange(1, npe + 1), dpn))
            and np.array_equal(ld, np.tile(dofs, npe))):
        return None
    nds = np.asarray(mod["nds"], dtype=np.float64)
    n_nodes = nds.shape[0]
    gl = np.asarray(gl_lookup)
    if n_dof != n_nodes * dpn or gl.shape[0] != n_nodes:
        return None
    expect = np.arange(n_nodes, dtype=np.int64)[:, None] * dpn + np.arange(dpn) + 1
    if gl.shape[1] < int(dofs.max()) or not np.array_equal(gl[:, dofs - 1], expect):
        return None
    if np.count_nonzero(gl) != n_nodes * dpn:
        return None
    els = np.asarray(mod["els"])[:, :npe]
    n_el = els.shape[0]
    mat = np.asarray(mod["el_mat_i"], dtype=np.int64).ravel()
    if mat.size != n_el or mat.min() < 1:
        return None
    el_nodes = (els - 1).astype(np.int32)
    # shape keys: node offsets from vertex 1, rounded; material in the key
    ext = 0.0
    for s0 in range(0, n_el, _CHUNK):
        e = el_nodes[s0:s0 + _CHUNK]
        ext = max(ext, float(np.abs(nds[e[:, 1:]] - nds[e[:, :1]]).max()))
    if not ext > 0:
        return None
    scale = 2.0 ** 24 / ext
    rng = np.random.default_rng(12345)
    mult = rng.integers(1, 2 ** 62, size=(npe - 1) * nds.shape[1] + 1, dtype=np.int64) | 1
    keys = np.empty(n_el, dtype=np.int64)
    qs = []
    for s0 in range(0, n_el, _CHUNK):
        e = el_nodes[s0:s0 + _CHUNK]
        q = np.rint((nds[e[:, 1:]] - nds[e[:, :1]]) * scale).astype(np.int64).reshape(e.shape[0], -1)
        q = np.concatenate([q, mat[s0:s0 + e.shape[0], None]], axis=1)
        with np.errstate(over="ignore"):
            keys[s0:s0 + e.shape[0]] = (q * mult).sum(axis=1)
        qs.append(q)
    uk, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    if uk.size * npe > _MAX_SLOTS:
        return None
    rep_q = np.concatenate([qs[f // _CHUNK][f % _CHUNK][None] for f in first])
    for k, s0 in enumerate(range(0, n_el, _CHUNK)):  # no hash collisions
        if not np.array_equal(qs[k], rep_q[inv[s0:s0 + qs[k].shape[0]]]):
            return None
    del qs
    from .assembly import _material
    tmpl = np.zeros((uk.size, npe * dpn, npe * dpn))
    for t, f in enumerate(first):
        D, rho = _material(matls, int(mat[f]))
        Ke = fn(nds, el_nodes[f:f + 1].astype(np.int64), D, rho, o["dof_to_use"])[0]
        tmpl[t] = Ke[0]
    a = np.asarray(mod.get("el_abs_i", np.zeros(n_el)), dtype=np.float64).ravel()
    p = float(o["damping_power_law"])
    el_scale = np.exp(np.log(float(o["max_stiffness_reduction"])) * a ** (p + 1))
    return ElementOperator(n_nodes=n_nodes, dpn=dpn, el_nodes=el_nodes,
                           el_tmpl=inv.astype(np.int32), el_scale=el_scale, tmpl=tmpl)


def ebe_bytes_per_dof(op: ElementOperator, n: int, damped: bool) -> float:
    """Algorithmic HBM bytes per DOF-step of the element-by-element step:
    per element its node indices (4 npe), template index (1) and scale (8);
    per DOF u_n and u_{n-1} read, u_{n+1} written (24) and the coefficient
    streams (alpha, plus beta when damped)."""
    npe = op.el_nodes.shape[1]
    return op.n_el / n * (4 * npe + 1 + 8) + 24 + 8 * (2 if damped else 1)


def ebe_stored_bytes_per_dof(op: ElementOperator, n: int, n_slot: int, damped: bool,
                             n_classes: int = 0) -> float:
    """Bytes per DOF-step this implementation streams if every record is read
    once: the slot-major node-to-element table (4 n_slot per node), and either
    (mesh-pattern kernel, n_classes > 0) the neighbour table (4 per neighbour
    class other than the node itself) and the element scale (8), or (generic
    kernel) the padded node quads (16) and the scale (8) per element; plus the
    vectors and coefficient streams."""
    if n_classes > 0:
        per_node, per_el = 4.0 * n_slot + 4.0 * (n_classes - 1), 8.0
    else:
        per_node, per_el = 4.0 * n_slot, 24.0
    return (per_node * op.n_nodes + per_el * op.n_el) / n + 24 + 8 * (2 if damped else 1)

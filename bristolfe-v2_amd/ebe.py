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
one step against the CSR operator before using it. A mesh of one element
type with every node carrying the same DOFs gives one group over all nodes;
a mesh of several types (solid CPE3 and fluid AC2D3 with ASI2D2 interface
elements, which have no stiffness) gives one group per type, each over its
own nodes with each node's first global DOF. Returns None when the model
does not have that form (a node's DOFs of one group not consecutive, too
many templates, unsupported element types).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .elements import _tri_geometry, element_function
from .options import set_default_fields

_NPE = {"CPE3": 3, "AC2D3": 3, "C3D4": 4}
_CHUNK = 1 << 20
_MAX_SLOTS = 64
DEVICE_MIN_ELEMENTS = 200_000  # template detection on the GPU above this many elements


@dataclass
class ElementOperator:
    """One element group: elements of one type over the group's own nodes
    0..n_nodes-1. node_dof[j] is the global DOF (0-based) of component 0 of
    node j; None means j*dpn (a single group over every node)."""
    n_nodes: int
    dpn: int
    el_nodes: np.ndarray   # (n_el, npe) int32, 0-based group-local nodes
    el_tmpl: np.ndarray    # (n_el,) int32
    el_scale: np.ndarray   # (n_el,) float64
    tmpl: np.ndarray       # (n_tmpl, npe*dpn, npe*dpn) float64
    node_dof: np.ndarray | None = None
    typ: str = ""

    @property
    def n_el(self) -> int:
        return int(self.el_nodes.shape[0])

    @property
    def groups(self):
        return [self]


@dataclass
class ElementGroups:
    """Element operator of a model whose nodes carry different DOF sets: one
    ElementOperator per element type, K = sum of the groups."""
    groups: list

    @property
    def n_el(self) -> int:
        return sum(g.n_el for g in self.groups)


def _unique_type(el_typ_i):
    a = np.asarray(el_typ_i, dtype=object).ravel()
    if a.size == 0 or not bool(np.all(a == a[0])):
        return None
    return a[0]


def _templates(nds, el_nodes, mat, matls, fn, dof_to_use, npe, dpn):
    """(el_tmpl, tmpl, factor) of a set of elements of one type, or None when
    there are too many distinct shapes. Shapes are node offsets from vertex 1
    rounded to 2^-24 of the largest element extent, with the material. Large
    sets are classified on the GPU (bfe_element_templates) when one is
    present, with the same keys, order and checks."""
    n_el = el_nodes.shape[0]
    rng = np.random.default_rng(12345)
    mult = rng.integers(1, 2 ** 62, size=(npe - 1) * nds.shape[1] + 1, dtype=np.int64) | 1
    found = None
    if n_el > DEVICE_MIN_ELEMENTS:
        found = _templates_device(nds, el_nodes, mat, mult, npe)
        if found is False:
            return None
    if found is None:
        found = _templates_host(nds, el_nodes, mat, mult, npe)
        if found is None:
            return None
    inv, first, factor = found
    from .assembly import _material
    tmpl = np.zeros((first.size, npe * dpn, npe * dpn))
    for t, f in enumerate(first):
        D, rho = _material(matls, int(mat[f]))
        tmpl[t] = fn(nds, el_nodes[f:f + 1].astype(np.int64), D, rho, dof_to_use)[0][0]
    return inv.astype(np.int32), tmpl, factor


def _templates_device(nds, el_nodes, mat, mult, npe):
    """(inv, first, factor) from the GPU, False when the set does not qualify
    (too many shapes, a zero extent, a hash collision), None without a GPU."""
    from . import _lib
    try:
        r = _lib.element_templates(nds, el_nodes, mat, mult, _MAX_SLOTS // npe,
                                   npe == 3 and nds.shape[1] == 2)
    except (OSError, RuntimeError) as e:
        if getattr(e, "status", None) == _lib.BFE_ERR_ARG:
            return False
        if getattr(e, "status", None) not in (None, _lib.BFE_ERR_HIP):
            raise
        return None
    if r is None:
        return False
    inv, first, jf = r
    return inv, first, (jf if jf is not None else np.ones(el_nodes.shape[0]))


def _templates_host(nds, el_nodes, mat, mult, npe):
    n_el = el_nodes.shape[0]
    ext = 0.0
    for s0 in range(0, n_el, _CHUNK):
        e = el_nodes[s0:s0 + _CHUNK]
        ext = max(ext, float(np.abs(nds[e[:, 1:]] - nds[e[:, :1]]).max()))
    if not ext > 0:
        return None
    scale = 2.0 ** 24 / ext
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
    # Triangle matrices are K_e = G(b, c) / J_e with J_e evaluated from absolute
    # coordinates (x1*y2 - y1*x2 - ..., fn_el_CPE3.m:56): far from the origin it
    # carries a cancellation error of about eps*|x|^2/h^2 that the edge vectors
    # b, c do not. Each element's J_rep / J_e goes into its scale, so the
    # operator keeps every element's own J and matches the assembled K to the
    # rounding of b and c.
    factor = np.ones(n_el)
    if npe == 3:
        Jrep = _tri_geometry(nds, el_nodes[first].astype(np.int64))[0]
        for s0 in range(0, n_el, _CHUNK):
            e = el_nodes[s0:s0 + _CHUNK].astype(np.int64)
            Je = _tri_geometry(nds, e)[0]
            factor[s0:s0 + e.shape[0]] = Jrep[inv[s0:s0 + e.shape[0]]] / Je
    return inv, first, factor


def element_operator(mod, matls, fe_options, gl_lookup, n_dof):
    """ElementOperator (one group over every node, node-major numbering) or
    ElementGroups (one group per element type with stiffness) for the model,
    or None if it does not qualify."""
    o = set_default_fields(dict(fe_options or {}), {
        "dof_to_use": [], "damping_power_law": 3, "max_stiffness_reduction": 0.01})
    els_all = np.asarray(mod["els"])
    n_all = els_all.shape[0]
    if "el_typ_i" in mod:
        typ_i = np.asarray(mod["el_typ_i"], dtype=object).ravel()
    else:
        typ_i = np.asarray([matls[i - 1]["el_typ"] for i in np.asarray(mod["el_mat_i"])], dtype=object)
    one = _unique_type(typ_i)
    types = [one] if one is not None else sorted(set(typ_i.tolist()))
    types = [t for t in types if t != "ASI2D2"]  # interface elements have no stiffness (C only)
    if not types or any(t not in _NPE for t in types):
        return None
    nds = np.asarray(mod["nds"], dtype=np.float64)
    n_nodes_all = nds.shape[0]
    gl = np.asarray(gl_lookup)
    mat_all = np.asarray(mod["el_mat_i"], dtype=np.int64).ravel()
    a_all = np.asarray(mod.get("el_abs_i", np.zeros(n_all)), dtype=np.float64).ravel()
    p = float(o["damping_power_law"])
    lmsr = np.log(float(o["max_stiffness_reduction"]))
    groups = []
    for typ in types:
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
        sel = np.nonzero(typ_i == typ)[0] if one is None else np.arange(n_all)
        els = els_all[sel, :npe] - 1
        mat = mat_all[sel]
        if mat.size and mat.min() < 1:
            return None
        if gl.shape[1] < int(dofs.max()):
            return None
        if len(types) == 1 and n_dof == n_nodes_all * dpn and gl.shape[0] == n_nodes_all:
            expect = np.arange(n_nodes_all, dtype=np.int64)[:, None] * dpn + np.arange(dpn) + 1
            if (np.array_equal(gl[:, dofs - 1], expect)
                    and np.count_nonzero(gl) == n_nodes_all * dpn):
                nodes, el_nodes, node_dof = None, els.astype(np.int32), None
            else:
                nodes = np.unique(els)
        else:
            nodes = np.unique(els)
        if nodes is not None:
            # the group's nodes in global order; a node's DOFs must be consecutive
            g0 = gl[nodes][:, dofs - 1].astype(np.int64)
            if np.any(g0 <= 0) or np.any(g0 - g0[:, :1] != np.arange(dpn)):
                return None
            node_dof = (g0[:, 0] - 1).astype(np.int32)
            loc = np.full(n_nodes_all, -1, dtype=np.int64)
            loc[nodes] = np.arange(nodes.size)
            el_nodes = loc[els].astype(np.int32)
        tm = _templates(nds, els.astype(np.int32), mat, matls, fn, o["dof_to_use"], npe, dpn)
        if tm is None:
            return None
        el_tmpl, tmpl, factor = tm
        el_scale = np.exp(lmsr * a_all[sel] ** (p + 1)) * factor
        groups.append(ElementOperator(n_nodes=int(n_nodes_all if nodes is None else nodes.size),
                                      dpn=dpn, el_nodes=el_nodes, el_tmpl=el_tmpl,
                                      el_scale=el_scale, tmpl=tmpl, node_dof=node_dof, typ=typ))
    if len(groups) == 1 and groups[0].node_dof is None:
        return groups[0]
    return ElementGroups(groups)


def ebe_bytes_per_dof(op, n: int, damped: bool) -> float:
    """Algorithmic HBM bytes per DOF-step of the element-by-element step:
    per element its node indices (4 npe), template index (1) and scale (8);
    per DOF u_n and u_{n-1} read, u_{n+1} written (24) and the coefficient
    streams (alpha, plus beta when damped)."""
    el = sum(g.n_el * (4 * g.el_nodes.shape[1] + 1 + 8) for g in op.groups)
    return el / n + 24 + 8 * (2 if damped else 1)

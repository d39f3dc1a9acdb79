"""Model set-up helpers (host, FP64): meshes, regions, absorbing layers,
fluid-solid interface elements, scatterers, materials, time steps, signals.

All node/element numbers are 1-based, as in the reference's ``mod`` struct:
``mod = {"nds": (n, d) float, "els": (m, k) int (0 = unused column),
"el_mat_i": (m,), "el_abs_i": (m,), ["el_typ_i": (m,) object]}``.

Reference counterparts (restated, vectorised):
  fn_rectangular_structured_mesh   code/fn_rectangular_structured_mesh.m:16-56
  fn_isometric_structured_mesh     code/fn_isometric_structured_mesh.m:16-84
  fn_calc_element_centres          code/fn_calc_element_centres.m:5-13
  fn_elements_in_region            code/fn_elements_in_region.m:5-8 (inpolygon)
  fn_remove_unused_nodes           code/internal/fn_remove_unused_nodes.m:1-18
  fn_dist_point_to_bdry_2D         code/fn_dist_point_to_bdry_2D.m:1-34
  fn_add_absorbing_layer           code/fn_add_absorbing_layer.m:10-17
  fn_get_edges                     code/fn_get_edges.m:1-35
  fn_add_fluid_solid_interface_els code/fn_add_fluid_solid_interface_els.m:1-74
  fn_set_els_inside_bdry_to_mat    code/fn_set_els_inside_bdry_to_mat.m:1-10
  fn_create_smooth_random_blob     code/fn_create_smooth_random_blob.m:1-14 (seeded)
  fn_find_nodes_on_line            code/fn_find_nodes_on_line.m:41-53
  fn_find_node_at_point            code/fn_find_node_at_point.m:15-21
  fn_isotropic_stiffness_matrix    code/fn_isotropic_stiffness_matrix.m:17-23
  fn_estimate_max_min_vels         code/fn_estimate_max_min_vels.m:1-30
  fn_get_suitable_time_step        code/fn_get_suitable_time_step.m:9-16
  fn_gaussian_pulse                code/fn_gaussian_pulse.m:1-15
New here: structured tetrahedral cube mesh, Hann toneburst, stable time step
from the largest eigenvalue of M^-1 K.

From 20k points or elements the heavy primitives run in libbfe_hip's
multi-threaded C++ (bfe_tri_grid, bfe_remove_unused_nodes, bfe_get_edges,
bfe_element_centres, bfe_points_in_polygon, bfe_dist_points_to_polygon),
with the same results as the numpy restatements kept beside them (_*_np).
The seeded blob (a few hundred points) stays in numpy.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

# ----------------------------------------------------------------------------
# meshes


def _model(nds, els):
    n = els.shape[0]
    return {"nds": nds, "els": els, "el_mat_i": np.ones(n, dtype=np.int64),
            "el_abs_i": np.zeros(n)}


def fn_rectangular_structured_mesh(bdry_pts, el_size):
    bdry_pts = np.asarray(bdry_pts, dtype=np.float64)
    lo, hi = bdry_pts.min(axis=0), bdry_pts.max(axis=0)
    nx = int(np.ceil(abs(hi[0] - lo[0]) / el_size)) + 1
    ny = int(np.ceil(abs(hi[1] - lo[1]) / el_size)) + 1
    x = np.linspace(lo[0], hi[0], nx)
    y = np.linspace(lo[1], hi[1], ny)
    lib = _native() if nx * ny >= _NATIVE_MIN else None
    if lib is not None:
        nds, els = lib.tri_grid(0, x, y, 0.0)
    else:
        nds, els = _rect_grid_np(x, y)
    mod = _model(nds, els)
    inside = fn_elements_in_region(mod, bdry_pts)
    mod["els"] = mod["els"][inside]
    mod["nds"], mod["els"] = fn_remove_unused_nodes(mod["nds"], mod["els"])[:2]
    return _model(mod["nds"], mod["els"])


def fn_isometric_structured_mesh(bdry_pts, el_size):
    bdry_pts = np.asarray(bdry_pts, dtype=np.float64)
    lo, hi = bdry_pts.min(axis=0), bdry_pts.max(axis=0)
    sin60 = np.sin(np.deg2rad(60.0))
    bx = abs(hi[0] - lo[0]) + el_size
    by = abs(hi[1] - lo[1])
    nx = int(np.ceil(bx / el_size)) + 1
    ny = int(np.ceil(by / el_size / sin60)) + 1
    x = np.linspace(lo[0] - el_size / 2, hi[0] + el_size / 2, nx)
    y = np.linspace(lo[1], hi[1], ny)
    lib = _native() if nx * ny >= _NATIVE_MIN else None
    if lib is not None:
        nds, els = lib.tri_grid(1, x, y, el_size / 4)
    else:
        nds, els = _iso_grid_np(x, y, el_size / 4)
    mod = _model(nds, els)
    inside = fn_elements_in_region(mod, bdry_pts)
    mod["els"] = mod["els"][inside]
    mod["nds"], mod["els"] = fn_remove_unused_nodes(mod["nds"], mod["els"])[:2]
    return _model(mod["nds"], mod["els"])


def _rect_grid_np(x, y):
    """Grid and triangles of code/fn_rectangular_structured_mesh.m:22-43
    before the clip (MATLAB meshgrid, column-major node numbers: node
    (iy, ix) -> iy + ny ix + 1)."""
    nx, ny = x.size, y.size
    X, Y = np.meshgrid(x, y)
    nds = np.stack([X.ravel(order="F"), Y.ravel(order="F")], axis=1)
    nn = np.arange(1, nx * ny + 1).reshape((ny, nx), order="F")
    a = np.stack([nn[:-1, :-1].ravel(order="F"), nn[1:, 1:].ravel(order="F"),
                  nn[1:, :-1].ravel(order="F")], axis=1)
    b = np.stack([nn[:-1, :-1].ravel(order="F"), nn[:-1, 1:].ravel(order="F"),
                  nn[1:, 1:].ravel(order="F")], axis=1)
    return nds, np.concatenate([a, b], axis=0)


def _iso_grid_np(x, y, shift):
    """Grid and triangles of code/fn_isometric_structured_mesh.m:22-75 before
    the clip (even rows moved by +shift, odd rows by -shift)."""
    nx, ny = x.size, y.size
    X, Y = np.meshgrid(x, y)
    X[0::2, :] += shift
    X[1::2, :] -= shift
    nn = np.arange(1, nx * ny + 1).reshape((ny, nx), order="F")
    f = lambda A: A.ravel(order="F")
    parts = [
        (nn[0:-1:2, :-1], nn[1::2, 1:], nn[1::2, :-1]),
        (nn[0:-1:2, :-1], nn[0:-1:2, 1:], nn[1::2, 1:]),
        (nn[1:-1:2, 1:], nn[2::2, 1:], nn[2::2, :-1]),
        (nn[1:-1:2, :-1], nn[1:-1:2, 1:], nn[2::2, :-1]),
    ]
    els = np.concatenate([np.stack([f(p), f(q), f(r)], axis=1) for p, q, r in parts], axis=0)
    return np.stack([f(X), f(Y)], axis=1), els


def cubic_structured_tet_mesh(size, el_size, origin=(0.0, 0.0, 0.0)):
    """Structured cube mesh, each cubic cell split into 6 positively oriented
    tetrahedra around its main diagonal. No counterpart in the reference
    (code/fn_3d_cubic_structured_mesh.m:35-44 emits 8-node hexes).
    Nodes are numbered x fastest, then y, then z."""
    size = np.broadcast_to(np.asarray(size, dtype=np.float64), (3,))
    nc = np.maximum(np.ceil(size / el_size).astype(np.int64), 1)
    axes = [np.linspace(origin[d], origin[d] + size[d], nc[d] + 1) for d in range(3)]
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    nds = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    nx, ny = nc[0] + 1, nc[1] + 1
    k, j, i = np.meshgrid(np.arange(nc[2]), np.arange(nc[1]), np.arange(nc[0]), indexing="ij")
    base = (i + nx * (j + ny * k)).ravel() + 1
    dx, dy, dz = 1, nx, nx * ny
    v = [base, base + dx, base + dx + dy, base + dy,
         base + dz, base + dx + dz, base + dx + dy + dz, base + dy + dz]
    # Kuhn split along v0 -> v6
    tets = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    # every cell has the same shape, so one cell fixes each tet type's
    # orientation: swap two vertices of the negatively oriented types
    corner = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                       [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)
    order = []
    for t in tets:
        x = corner[list(t)]
        if np.linalg.det(x[1:] - x[:1]) < 0:
            t = (t[0], t[2], t[1], t[3])
        order.append(t)
    els = np.empty((len(tets) * base.size, 4), dtype=np.int64)
    for k, t in enumerate(order):
        for j, a in enumerate(t):
            els[k * base.size:(k + 1) * base.size, j] = v[a]
    return _model(nds, els)


# ----------------------------------------------------------------------------
# regions and geometry


_NATIVE_MIN = 20_000  # points or elements from which the C++ helpers are used


def _native():
    """libbfe_hip's host helpers, or None when the library is not built."""
    try:
        from . import _lib
        return _lib if _lib.load() is not None else None
    except (OSError, RuntimeError):
        return None


def inpolygon(px, py, vx, vy):
    """Points inside or on the edge of a closed polygon (MATLAB inpolygon)."""
    px = np.asarray(px, dtype=np.float64)
    lib = _native() if px.size >= _NATIVE_MIN else None
    if lib is not None:
        pts = np.stack([px.ravel(), np.asarray(py, dtype=np.float64).ravel()], axis=1)
        return lib.points_in_polygon(pts, vx, vy).reshape(px.shape)
    return _inpolygon_np(px, py, vx, vy)


def _inpolygon_np(px, py, vx, vy):
    px = np.asarray(px, dtype=np.float64)
    py = np.asarray(py, dtype=np.float64)
    vx = np.asarray(vx, dtype=np.float64).ravel()
    vy = np.asarray(vy, dtype=np.float64).ravel()
    res = np.zeros(px.shape, dtype=bool)
    span = max(vx.max() - vx.min(), vy.max() - vy.min(), 1e-300)
    eps = 1e-12 * span
    cand = np.nonzero((px >= vx.min() - eps) & (px <= vx.max() + eps) &
                      (py >= vy.min() - eps) & (py <= vy.max() + eps))
    x = px[cand]
    y = py[cand]
    inside = np.zeros(x.shape, dtype=bool)
    on = np.zeros(x.shape, dtype=bool)
    nv = vx.size
    for k in range(nv):
        x1, y1 = vx[k], vy[k]
        x2, y2 = vx[(k + 1) % nv], vy[(k + 1) % nv]
        cond = (y1 > y) != (y2 > y)
        if y2 != y1:
            xint = x1 + (y - y1) * ((x2 - x1) / (y2 - y1))
            inside ^= cond & (x < xint)
        cross = (x2 - x1) * (y - y1) - (y2 - y1) * (x - x1)
        scale = max(abs(x2 - x1), abs(y2 - y1), 1e-300)
        within = ((x >= min(x1, x2) - 1e-12 * scale) & (x <= max(x1, x2) + 1e-12 * scale) &
                  (y >= min(y1, y2) - 1e-12 * scale) & (y <= max(y1, y2) + 1e-12 * scale))
        on |= within & (np.abs(cross) <= 1e-12 * scale * scale)
    res[cand] = inside | on
    return res


def fn_calc_element_centres(nds, els):
    els = np.asarray(els)
    lib = _native() if els.shape[0] >= _NATIVE_MIN else None
    if lib is not None:
        return lib.element_centres(nds, els)
    return _element_centres_np(nds, els)


def _element_centres_np(nds, els):
    els = np.asarray(els)
    nds = np.asarray(nds)
    out = np.empty((els.shape[0], nds.shape[1]))
    for s0 in range(0, els.shape[0], 1 << 20):  # bounded temporaries on large meshes
        e = els[s0:s0 + (1 << 20)]
        valid = e > 0
        pts = nds[np.where(valid, e - 1, 0)] * valid[:, :, None]
        out[s0:s0 + e.shape[0]] = pts.sum(axis=1) / valid.sum(axis=1)[:, None]
    return out


def fn_elements_in_region(mod, region):
    c = fn_calc_element_centres(mod["nds"], mod["els"])
    region = np.asarray(region, dtype=np.float64)
    lib = _native() if c.shape[0] >= _NATIVE_MIN else None
    if lib is not None and c.shape[1] == 2:
        return lib.points_in_polygon(c, region[:, 0], region[:, 1])
    return inpolygon(c[:, 0], c[:, 1], region[:, 0], region[:, 1])


def fn_remove_unused_nodes(nds, els):
    els = np.asarray(els)
    n = np.asarray(nds).shape[0]
    lib = _native() if els.size >= _NATIVE_MIN else None
    if lib is not None:
        els2, old_nds = lib.remove_unused_nodes(n, els)
        new_nds = np.zeros(n, dtype=np.int64)
        new_nds[old_nds - 1] = np.arange(1, old_nds.size + 1)
        return np.asarray(nds)[old_nds - 1], els2.reshape(els.shape), old_nds, new_nds
    return _remove_unused_nodes_np(nds, els)


def _remove_unused_nodes_np(nds, els):
    els = np.asarray(els)
    in_use = np.zeros(np.asarray(nds).shape[0], dtype=bool)
    u = els[els > 0]
    in_use[u - 1] = True
    old_nds = np.nonzero(in_use)[0] + 1
    new_nds = np.zeros(in_use.size, dtype=np.int64)
    new_nds[old_nds - 1] = np.arange(1, old_nds.size + 1)
    els2 = np.where(els > 0, new_nds[np.maximum(els, 1) - 1], 0)
    return np.asarray(nds)[in_use], els2, old_nds, new_nds


def fn_remove_unused_elements(in_use, *arrays):
    in_use = np.asarray(in_use, dtype=bool)
    return tuple(np.asarray(a)[in_use] for a in arrays)


def fn_dist_point_to_bdry_2D(pts, bdry):
    """Distance from each point to a polygon, as the reference computes it:
    the nearest vertex and the two edges adjacent to it (fn_dist_point_to_bdry_2D.m:6-22)."""
    pts = np.asarray(pts, dtype=np.float64)
    lib = _native() if pts.shape[0] >= _NATIVE_MIN else None
    if lib is not None:
        return lib.dist_points_to_polygon(pts, bdry)
    return _dist_point_to_bdry_2D_np(pts, bdry)


def _dist_point_to_bdry_2D_np(pts, bdry):
    pts = np.asarray(pts, dtype=np.float64)
    bdry = np.asarray(bdry, dtype=np.float64)
    nb = bdry.shape[0]
    r0all = np.sqrt(((pts[:, None, :] - bdry[None, :, :]) ** 2).sum(axis=2))
    j = np.argmin(r0all, axis=1)
    r0 = r0all[np.arange(pts.shape[0]), j]
    j1 = (j + 1) % nb
    j2 = (j - 1) % nb

    def dist_line(p1, p2):
        n = p2 - p1
        d = np.sqrt((n ** 2).sum(axis=1))
        n = n / d[:, None]
        s = ((pts - p1) * n).sum(axis=1)
        r = pts - p1 - s[:, None] * n
        s = s / d
        r = np.sqrt((r ** 2).sum(axis=1))
        return np.where((s < 0) | (s > 1), np.inf, r)

    r1 = dist_line(bdry[j], bdry[j1])
    r2 = dist_line(bdry[j], bdry[j2])
    return np.minimum(np.minimum(r0, r1), r2)


def fn_add_absorbing_layer(mod, abs_bdry_pts, abs_bdry_thickness):
    c = fn_calc_element_centres(mod["nds"], mod["els"])
    d = fn_dist_point_to_bdry_2D(c, abs_bdry_pts)
    inside = fn_elements_in_region(mod, abs_bdry_pts)
    a = np.clip(d / abs_bdry_thickness, 0.0, 1.0)
    a[inside] = 0.0
    mod = dict(mod)
    mod["el_abs_i"] = a
    return mod


def add_absorbing_layer_box(mod, inner_lo, inner_hi, thickness):
    """3D form of fn_add_absorbing_layer (code/fn_add_absorbing_layer.m:10-17)
    for a box-shaped inner region: a_e = clip(d_e / thickness, 0, 1) with d_e
    the distance from the element centroid to the box [inner_lo, inner_hi]
    (0 inside). An infinite bound leaves that axis without a layer, e.g.
    (x0, y0, -inf)..(x1, y1, inf) puts layers on the four side faces only."""
    c = fn_calc_element_centres(mod["nds"], mod["els"])
    lo = np.asarray(inner_lo, dtype=np.float64)
    hi = np.asarray(inner_hi, dtype=np.float64)
    d2 = np.zeros(c.shape[0])
    for k in range(c.shape[1]):
        g = np.maximum(np.maximum(lo[k] - c[:, k], c[:, k] - hi[k]), 0.0)
        d2 += g * g
    mod = dict(mod)
    mod["el_abs_i"] = np.clip(np.sqrt(d2) / thickness, 0.0, 1.0)
    return mod


def fn_get_edges(els):
    """Unique edges (sorted node pairs, sorted rows) and the elements on each side
    (largest element number first, 0 where there is none)."""
    els = np.asarray(els)
    lib = _native() if els.shape[0] >= _NATIVE_MIN else None
    if lib is not None:
        ed, pt, el = lib.get_edges(els)
        cnt = np.diff(pt)
        width = int(cnt.max()) if cnt.size else 0
        ewe = np.zeros((ed.shape[0], width), dtype=np.int64)
        row = np.repeat(np.arange(ed.shape[0]), cnt)
        ewe[row, np.arange(el.size) - pt[row]] = el
        return ed, ewe
    return _get_edges_np(els)


def _get_edges_np(els):
    els = np.asarray(els)
    eds, owner = [], []
    nper = (els > 0).sum(axis=1)
    for k in np.unique(nper):
        sel = np.nonzero(nper == k)[0]
        e = els[sel, :k]
        for a in range(k):
            b = (a + 1) % k
            eds.append(np.stack([e[:, a], e[:, b]], axis=1))
            owner.append(sel + 1)
    ed = np.sort(np.concatenate(eds), axis=1)
    own = np.concatenate(owner)
    order = np.lexsort((ed[:, 1], ed[:, 0]))
    ed, own = ed[order], own[order]
    uniq, inv = np.unique(ed, axis=0, return_inverse=True)
    inv = inv.ravel()
    cnt = np.bincount(inv)
    width = cnt.max() if cnt.size else 0
    els_with_ed = np.zeros((uniq.shape[0], width), dtype=np.int64)
    # descending element numbers per edge
    o2 = np.lexsort((-own, inv))
    inv_s, own_s = inv[o2], own[o2]
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    pos = np.arange(inv_s.size) - start[inv_s]
    els_with_ed[inv_s, pos] = own_s
    return uniq, els_with_ed


def fn_find_free_edges(els):
    """Edges with exactly one element on them and two nodes
    (code/fn_find_free_edges.m:8-10), sorted as fn_get_edges sorts them."""
    ed, ewe = fn_get_edges(els)
    j = ((ewe > 0).sum(axis=1) == 1) & ((ed > 0).sum(axis=1) == 2)
    return ed[j]


def fn_add_scatterer(mod, matls, scat_pts, scat_matl):
    """code/fn_add_scatterer.m:1-37: removes the interface elements, then sets
    the elements whose centres lie inside scat_pts to material scat_matl, or,
    for scat_matl = 0, removes them and the nodes left unused (a subdomain
    model's bdry_lyrs and main_nd_i follow its nodes), and adds the
    fluid-solid interface elements again; int_el_i flags the elements inside
    the subdomain's inner boundary."""
    mod = dict(mod)
    keep = np.asarray(mod["el_typ_i"], dtype=object) != "ASI2D2"
    for k in ("els", "el_mat_i", "el_abs_i", "el_typ_i"):
        mod[k] = np.asarray(mod[k])[keep]
    if scat_matl > 0:
        mod.pop("el_typ_i")
        mod = fn_set_els_inside_bdry_to_mat(mod, scat_pts, scat_matl)
    else:
        keep = ~fn_elements_in_region(mod, scat_pts)
        for k in ("els", "el_mat_i", "el_abs_i", "el_typ_i"):
            mod[k] = np.asarray(mod[k])[keep]
        mod["nds"], mod["els"], old_nds, _ = fn_remove_unused_nodes(mod["nds"], mod["els"])
        for k in ("bdry_lyrs", "main_nd_i"):
            if k in mod:
                mod[k] = np.asarray(mod[k])[old_nds - 1]
    mod = fn_add_fluid_solid_interface_els(mod, matls)
    if "inner_bndry_pts" in mod:
        mod["int_el_i"] = fn_elements_in_region(mod, mod["inner_bndry_pts"])
    return mod


def fn_add_fluid_solid_interface_els(mod, matls, options=None):
    opts = {"interface_el_name": "ASI2D2", "fluid_el_names": ["AC2D3"], "solid_el_names": ["CPE3"]}
    opts.update(options or {})
    mod = dict(mod)
    if "el_typ_i" not in mod:
        mod["el_typ_i"] = np.asarray([matls[i - 1]["el_typ"] for i in mod["el_mat_i"]], dtype=object)
    typ = np.asarray(mod["el_typ_i"], dtype=object)
    els = np.asarray(mod["els"])
    # Only elements touching a node shared by fluid and solid elements can own
    # an interface edge; restricting the edge search to them gives the same
    # edges in the same (sorted) order.
    is_fl = np.isin(typ, opts["fluid_el_names"])
    is_so = np.isin(typ, opts["solid_el_names"])
    nn = np.asarray(mod["nds"]).shape[0]
    in_fl = np.zeros(nn + 1, dtype=bool)
    in_so = np.zeros(nn + 1, dtype=bool)
    in_fl[els[is_fl].ravel()] = True
    in_so[els[is_so].ravel()] = True
    shared = in_fl & in_so
    shared[0] = False
    cand = np.nonzero(np.any(shared[els], axis=1) & (is_fl | is_so))[0]
    ed, ewe_local = fn_get_edges(els[cand]) if cand.size else (np.zeros((0, 2), np.int64),
                                                              np.zeros((0, 2), np.int64))
    ewe = np.where(ewe_local > 0, cand[np.maximum(ewe_local, 1) - 1] + 1, 0)
    if ewe.shape[1] < 2:
        ewe = np.concatenate([ewe, np.zeros((ewe.shape[0], 2 - ewe.shape[1]), np.int64)], axis=1)
    t = np.where(ewe > 0, typ[np.maximum(ewe, 1) - 1], "")
    isf = np.isin(t, opts["fluid_el_names"])
    iss = np.isin(t, opts["solid_el_names"])
    j = (isf[:, 0] & iss[:, 1]) | (isf[:, 1] & iss[:, 0])
    ed, ewe, isf = ed[j].copy(), ewe[j], isf[j]
    nds = np.asarray(mod["nds"])
    fe = np.where(isf[:, 0], ewe[:, 0], ewe[:, 1])
    ec = fn_calc_element_centres(nds, els[fe - 1]) if fe.size else np.zeros((0, 2))
    a = nds[ed[:, 1] - 1] - nds[ed[:, 0] - 1]
    b = np.stack([a[:, 1], -a[:, 0]], axis=1)
    c = ec - nds[ed[:, 0] - 1]
    flip = (c * b).sum(axis=1) < 0
    ed[flip] = ed[flip][:, ::-1]
    n_int = ed.shape[0]
    pad = np.zeros((n_int, els.shape[1]), dtype=els.dtype)
    pad[:, :2] = ed
    mod["els"] = np.concatenate([els, pad], axis=0)
    mod["el_typ_i"] = np.concatenate([typ, np.asarray([opts["interface_el_name"]] * n_int, dtype=object)])
    mod["el_mat_i"] = np.concatenate([np.asarray(mod["el_mat_i"]), np.zeros(n_int, np.int64)])
    if "el_abs_i" in mod:
        mod["el_abs_i"] = np.concatenate([np.asarray(mod["el_abs_i"]), np.zeros(n_int)])
    return mod


def fn_set_els_inside_bdry_to_mat(mod, bdry_pts, mat_i):
    mod = dict(mod)
    inside = fn_elements_in_region(mod, bdry_pts)
    mi = np.asarray(mod["el_mat_i"]).copy()
    mi[inside] = mat_i
    mod["el_mat_i"] = mi
    return mod


def fn_create_smooth_random_blob(min_rad_frac, complexity, no_pts, rng=None):
    """Seeded restatement: phases drawn from ``rng`` (numpy Generator)."""
    rng = np.random.default_rng(rng)
    a = np.linspace(0.0, 2.0 * np.pi, no_pts + 1)[:-1][:, None]
    if complexity > 0:
        n = np.arange(complexity + 1)[None, :]
        phi = rng.random(n.shape) * 2.0 * np.pi
        r = np.sin(a * n + phi).sum(axis=1)
        r = (r - r.min()) / (r.max() - r.min()) * (1 - min_rad_frac) + min_rad_frac
    else:
        r = np.ones(no_pts)
    return r[:, None] * np.concatenate([np.cos(a), np.sin(a)], axis=1)


def remove_elements_in_region(mod, region_pts, matls=None, add_interfaces=False):
    """Void scatterer: drop elements whose centre lies in ``region_pts``, then
    unused nodes (code/fn_add_scatterer.m:14-23 with scat_matl = 0)."""
    mod = dict(mod)
    typ = np.asarray(mod.get("el_typ_i", np.array(["" for _ in range(len(mod["el_mat_i"]))], dtype=object)))
    keep = ~fn_elements_in_region(mod, region_pts)
    if "el_typ_i" in mod:
        keep &= typ != "ASI2D2"
    for key in ("els", "el_mat_i", "el_abs_i", "el_typ_i"):
        if key in mod:
            mod[key] = np.asarray(mod[key])[keep]
    mod["nds"], mod["els"], _, _ = fn_remove_unused_nodes(mod["nds"], mod["els"])
    if add_interfaces:
        mod = fn_add_fluid_solid_interface_els(mod, matls)
    return mod


def fn_find_nodes_on_line(nodes, p1, p2, tol):
    nodes = np.asarray(nodes, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    p2 = np.asarray(p2, dtype=np.float64)
    n = p2 - p1
    d = np.sqrt((n ** 2).sum())
    n = n / d
    s = ((nodes - p1) * n).sum(axis=1)
    r = nodes - p1 - s[:, None] * n
    s = s / d
    r = np.sqrt((r ** 2).sum(axis=1))
    return np.nonzero((r <= tol) & (s >= -tol) & (s <= 1 + tol))[0] + 1


def fn_find_node_at_point(nodes, p, tol=np.inf):
    """Nearest node (first one on ties, as MATLAB min); 0 if farther than tol."""
    nodes = np.asarray(nodes, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    # search a growing window first: the nearest node inside a window of
    # half-width w is the global nearest when its distance is <= w
    span = float(np.max(nodes.max(axis=0) - nodes.min(axis=0))) if nodes.size else 0.0
    w = span / max(np.sqrt(nodes.shape[0]), 1.0) * 4.0
    while 0 < w < 2 * span:
        sel = np.nonzero(np.all(np.abs(nodes - p) <= w, axis=1))[0]
        if sel.size:
            r = np.sqrt(((nodes[sel] - p) ** 2).sum(axis=1))
            k = int(np.argmin(r))
            if r[k] <= w:
                i = int(sel[k])
                return 0 if r[k] > tol else i + 1
        w *= 2.0
    r = np.sqrt(((nodes - p) ** 2).sum(axis=1))
    i = int(np.argmin(r))
    return 0 if r[i] > tol else i + 1


def find_nodes_at_points(nodes, pts, tol=np.inf):
    """Vectorised fn_find_node_at_point for many points (same tie rule:
    the lowest node number among equidistant nearest nodes)."""
    from scipy.spatial import cKDTree

    nodes = np.asarray(nodes, dtype=np.float64)
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    k = min(16, nodes.shape[0])
    d, idx = cKDTree(nodes).query(pts, k=k)
    d = np.atleast_2d(d).reshape(pts.shape[0], k)
    idx = np.atleast_2d(idx).reshape(pts.shape[0], k)
    out = np.zeros(pts.shape[0], dtype=np.int64)
    for i in range(pts.shape[0]):
        r = np.sqrt(((nodes[idx[i]] - pts[i]) ** 2).sum(axis=1))  # exact distances
        best = r.min()
        j = idx[i][r == best].min()
        out[i] = 0 if best > tol else j + 1
    return out


# ----------------------------------------------------------------------------
# materials, time step, signals


def fn_isotropic_stiffness_matrix(E, v):
    c = E / (1 - 2 * v) / (1 + v)
    D = np.array([
        [1 - v, v, v, 0, 0, 0],
        [v, 1 - v, v, 0, 0, 0],
        [v, v, 1 - v, 0, 0, 0],
        [0, 0, 0, (1 - 2 * v) / 2, 0, 0],
        [0, 0, 0, 0, (1 - 2 * v) / 2, 0],
        [0, 0, 0, 0, 0, (1 - 2 * v) / 2],
    ], dtype=np.float64)
    return c * D


def fn_estimate_max_min_vels(matls):
    v = []
    for m in matls:
        s = np.asarray(m.get("D", m.get("stiffness_matrix", [])), dtype=np.float64).ravel()
        rho = m["density"] if "density" in m else m["rho"]
        s = s[np.abs(s) > 0]
        if s.size:
            v.append((np.sqrt(s.min() / rho), np.sqrt(s.max() / rho)))
    v = np.asarray(v)
    return v[:, 1].max(), v[:, 0].min()


def fn_get_suitable_time_step(matls, el_size, safety_factor=np.sqrt(2.0)):
    mx, _ = fn_estimate_max_min_vels(matls)
    return el_size / mx / safety_factor


def max_eigenvalue_MinvK(K, m, iters=300, tol=1e-10, seed=0):
    """Largest eigenvalue of M^-1 K by power iteration on the symmetric
    similar matrix |M|^-1/2 K |M|^-1/2 sign(M) (K and M share signs per DOF
    for every element type: solids positive, fluid pressure negative)."""
    m = np.asarray(m, dtype=np.float64)
    s = 1.0 / np.sqrt(np.abs(m))
    A = sp.diags(s) @ (sp.diags(np.sign(m)) @ K.tocsr()) @ sp.diags(s)
    A = A.tocsr()
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(K.shape[0])
    x /= np.linalg.norm(x)
    lam = 0.0
    for _ in range(iters):
        y = A @ x
        lam_new = float(x @ y)
        nrm = np.linalg.norm(y)
        if nrm == 0:
            return 0.0
        x = y / nrm
        if abs(lam_new - lam) <= tol * abs(lam_new):
            lam = lam_new
            break
        lam = lam_new
    return lam


def stable_time_step(K, m, safety=0.9, **kw):
    """dt = safety * 2 / omega_max with omega_max^2 the largest eigenvalue of
    M^-1 K, the stability limit of the undamped central-difference march."""
    lam = max_eigenvalue_MinvK(K, m, **kw)
    return safety * 2.0 / np.sqrt(lam)


def fn_gaussian_pulse(t, centre_freq, no_cycles):
    t = np.asarray(t, dtype=np.float64)
    sz = t.shape
    t = t.ravel()
    beta = 0.01
    ct = t.mean()
    T = no_cycles / (2 * centre_freq * np.sqrt(np.log(1 / beta)))
    env = np.exp(-((t - ct) / T) ** 2)
    i = int(np.nonzero(env > 0.01)[0].min())
    s = env * np.sin(2 * np.pi * centre_freq * (t - ct))
    s = np.concatenate([s[i:], np.zeros(i)])
    return s.reshape(sz)


def hann_toneburst(t, centre_freq, no_cycles):
    """0.5 (1 - cos(2 pi f t / N)) sin(2 pi f t) on [0, N/f], zero elsewhere."""
    t = np.asarray(t, dtype=np.float64)
    dur = no_cycles / centre_freq
    w = 0.5 * (1.0 - np.cos(2 * np.pi * centre_freq * t / no_cycles))
    s = w * np.sin(2 * np.pi * centre_freq * t)
    return np.where((t >= 0) & (t <= dur), s, 0.0)

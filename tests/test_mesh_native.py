"""Host C++ mesh helpers (bfe_element_centres, bfe_points_in_polygon,
bfe_dist_points_to_polygon, bfe_tri_grid, bfe_remove_unused_nodes,
bfe_get_edges) against the numpy restatements of
code/fn_calc_element_centres.m, inpolygon, fn_dist_point_to_bdry_2D.m, the
structured meshers, fn_remove_unused_nodes.m and fn_get_edges.m: identical
to the last bit."""
import numpy as np

from conftest import PKG, coupled_block


def _lib():
    from bristolfe_v2_amd import _lib
    return _lib


def test_element_centres_bit_identical():
    mod, _ = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3)   # triangles and 2-node ASI2D2 rows
    els = np.asarray(mod["els"])
    assert (els == 0).any()
    a = _lib().element_centres(mod["nds"], els)
    b = PKG.mesh._element_centres_np(mod["nds"], els)
    np.testing.assert_array_equal(a, b)
    tets = PKG.mesh.cubic_structured_tet_mesh(1e-3, 0.25e-3)
    np.testing.assert_array_equal(_lib().element_centres(tets["nds"], tets["els"]),
                                  PKG.mesh._element_centres_np(tets["nds"], tets["els"]))


def test_points_in_polygon_bit_identical_incl_edges_and_vertices():
    rng = np.random.default_rng(0)
    blob = PKG.mesh.fn_create_smooth_random_blob(0.5, 3, 48, 9) * 1e-3 + np.array([2e-3, 1e-3])
    vx, vy = blob[:, 0], blob[:, 1]
    pts = np.vstack([rng.uniform([0.5e-3, -0.5e-3], [3.5e-3, 2.5e-3], (50_000, 2)),
                     blob,                                         # vertices
                     (blob + np.roll(blob, -1, axis=0)) / 2])      # edge midpoints
    a = _lib().points_in_polygon(pts, vx, vy)
    b = PKG.mesh._inpolygon_np(pts[:, 0], pts[:, 1], vx, vy)
    np.testing.assert_array_equal(a, b)
    assert a[-2 * len(blob):].all() and 0 < a.sum() < len(a)
    # a rectangle and points on its sides
    r = np.array([[0, 0], [2, 0], [2, 1], [0, 1]], dtype=float)
    p = np.array([[1, 0], [2, 0.5], [1, 1], [0, 0.5], [1, 0.5], [3, 0.5], [1, -1e-9]])
    np.testing.assert_array_equal(_lib().points_in_polygon(p, r[:, 0], r[:, 1]),
                                  [True, True, True, True, True, False, False])


def test_dist_points_to_polygon_bit_identical():
    rng = np.random.default_rng(1)
    inner = np.array([[0.6e-3, 0.6e-3], [2.4e-3, 0.6e-3], [2.4e-3, 5e-3], [0.6e-3, 5e-3]])
    pts = rng.uniform(0, 3e-3, (40_000, 2))
    a = _lib().dist_points_to_polygon(pts, inner)
    b = PKG.mesh._dist_point_to_bdry_2D_np(pts, inner)
    np.testing.assert_array_equal(a, b)
    # closed forms: left of the left side, below the bottom side
    q = np.array([[0.1e-3, 1e-3], [1e-3, 0.2e-3]])
    np.testing.assert_allclose(_lib().dist_points_to_polygon(q, inner), [0.5e-3, 0.4e-3],
                               rtol=1e-12)


def test_absorbing_layer_uses_native_helpers_identically():
    """fn_add_absorbing_layer on a mesh above the native threshold equals
    the all-numpy computation."""
    m = PKG.mesh
    mod = m.fn_rectangular_structured_mesh([[0, 0], [20e-3, 0], [20e-3, 10e-3], [0, 10e-3]], 0.05e-3)
    assert mod["els"].shape[0] >= m._NATIVE_MIN
    inner = np.array([[2e-3, 2e-3], [18e-3, 2e-3], [18e-3, 50e-3], [2e-3, 50e-3]])
    a = m.fn_add_absorbing_layer(mod, inner, 2e-3)["el_abs_i"]
    c = m._element_centres_np(mod["nds"], mod["els"])
    d = m._dist_point_to_bdry_2D_np(c, inner)
    inside = m._inpolygon_np(c[:, 0], c[:, 1], inner[:, 0], inner[:, 1])
    b = np.clip(d / 2e-3, 0.0, 1.0)
    b[inside] = 0.0
    np.testing.assert_array_equal(a, b)


def test_structured_grids_identical_to_numpy():
    """bfe_tri_grid against the numpy grids of fn_rectangular_structured_mesh
    and fn_isometric_structured_mesh (odd and even row counts, the 2 x 2
    minimum), and the meshers above the native threshold against the
    all-numpy path (region clip and node removal included)."""
    m = PKG.mesh
    for iso in (0, 1):
        for nx, ny in ((2, 2), (3, 2), (5, 4), (6, 7), (301, 150)):
            x = np.linspace(-1e-3, 2e-3, nx)
            y = np.linspace(0.5e-3, 3e-3, ny)
            a = _lib().tri_grid(iso, x, y, 0.25e-3 if iso else 0.0)
            b = m._iso_grid_np(x, y, 0.25e-3) if iso else m._rect_grid_np(x, y)
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])
    # a clipped shape above the threshold: native grid, clip and renumbering
    bdry = np.array([[0, 0], [8e-3, 0], [8e-3, 3e-3], [4e-3, 6e-3], [0, 6e-3]])
    for fn in (m.fn_rectangular_structured_mesh, m.fn_isometric_structured_mesh):
        nat = fn(bdry, 0.04e-3)
        assert nat["els"].shape[0] > m._NATIVE_MIN
        saved = m._native
        try:
            m._native = lambda: None
            ref = fn(bdry, 0.04e-3)
        finally:
            m._native = saved
        np.testing.assert_array_equal(nat["nds"], ref["nds"])
        np.testing.assert_array_equal(nat["els"], ref["els"])


def test_remove_unused_nodes_identical_to_numpy():
    rng = np.random.default_rng(5)
    n = 50_000
    els = rng.integers(1, n + 1, size=(30_000, 3))
    els[rng.random(els.shape) < 0.1] = 0            # absent columns
    nds = rng.standard_normal((n, 2))
    a = PKG.mesh.fn_remove_unused_nodes(nds, els)
    b = PKG.mesh._remove_unused_nodes_np(nds, els)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert a[2].size < n                              # some nodes were unused


def test_get_edges_identical_to_numpy():
    """bfe_get_edges against the numpy fn_get_edges: a triangle mesh, the
    coupled model with its 2-node interface rows (an element of k nodes
    gives k edges, as code/fn_get_edges.m:11-17 builds them), and tets."""
    m = PKG.mesh
    mod, _ = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.02e-3)
    tets = m.cubic_structured_tet_mesh(1e-3, 0.05e-3)
    tri = m.fn_rectangular_structured_mesh(np.array([[0, 0], [4e-3, 0], [4e-3, 2e-3], [0, 2e-3]]),
                                           0.02e-3)
    for els in (np.asarray(mod["els"]), tets["els"], tri["els"]):
        assert els.shape[0] >= m._NATIVE_MIN
        a = m.fn_get_edges(els)
        b = m._get_edges_np(els)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
    # a closed form: two triangles sharing an edge
    ed, pt, el = _lib().get_edges(np.array([[1, 2, 3], [2, 4, 3]]))
    np.testing.assert_array_equal(ed, [[1, 2], [1, 3], [2, 3], [2, 4], [3, 4]])
    np.testing.assert_array_equal(el[pt[2]:pt[3]], [2, 1])
    # interface elements found with the native edge search equal the numpy ones
    big, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.05e-3)
    keep = np.asarray(big["el_typ_i"], dtype=object) != "ASI2D2"
    base = {k: (np.asarray(v)[keep] if k in ("els", "el_mat_i", "el_abs_i", "el_typ_i") else v)
            for k, v in big.items()}
    saved, saved_min = m._native, m._NATIVE_MIN
    try:
        m._NATIVE_MIN = 1           # every helper native, the edge search included
        nat = m.fn_add_fluid_solid_interface_els(base, matls)
        m._NATIVE_MIN = saved_min
        m._native = lambda: None    # every helper numpy
        ref = m.fn_add_fluid_solid_interface_els(base, matls)
    finally:
        m._native, m._NATIVE_MIN = saved, saved_min
    np.testing.assert_array_equal(nat["els"], ref["els"])
    np.testing.assert_array_equal(np.asarray(nat["el_typ_i"]), np.asarray(ref["el_typ_i"]))

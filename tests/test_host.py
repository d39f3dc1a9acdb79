"""Host-side (CPU) logic: assembly against the oracle's per-element loop,
mesh and model helpers, driver argument mapping, field output, options."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import PKG, assemble, coupled_block, rect, rel_l2, solid_plate, steel, water
from oracle import assembly as oa
from oracle import post as op


def _cmp_assembly(mod, matls, opts):
    K, C, M, gl = assemble(mod, matls, opts)
    K2, C2, M2, gl2 = oa.build_global_matrices(mod["nds"], mod["els"], mod["el_mat_i"], mod["el_abs_i"],
                                               list(mod["el_typ_i"]), matls, opts)
    assert K.shape == K2.shape
    np.testing.assert_array_equal(gl, gl2)
    for A, B in ((K, K2), (C, C2), (M, M2)):
        scale = max(abs(B).max(), 1e-300)
        assert abs(A - B).max() <= 1e-13 * scale
    return K, C, M, gl


@pytest.mark.parametrize("opts", [{"dof_to_use": [1, 2]}, {}, {"dof_to_use": [1, 2, 4]}])
def test_assembly_coupled_matches_oracle(opts):
    mod, matls = coupled_block(2e-3, 1e-3, 0.75e-3, 0.25e-3)
    _cmp_assembly(mod, matls, dict(opts, damping_power_law=2))


def test_assembly_void_and_isometric_mesh():
    m = PKG.mesh
    matls = [steel(), water()]
    mod = m.fn_isometric_structured_mesh(rect(2e-3, 2e-3), 0.2e-3)
    mod = m.fn_set_els_inside_bdry_to_mat(mod, rect(3e-3, 1e-3, -1e-3, 1.2e-3), 2)
    mod["el_typ_i"] = np.asarray([matls[i - 1]["el_typ"] for i in mod["el_mat_i"]], dtype=object)
    blob = m.fn_create_smooth_random_blob(0.5, 3, 32, 3) * 0.35e-3 + np.array([1e-3, 0.6e-3])
    mod = m.remove_elements_in_region(mod, blob)
    mod = m.fn_add_fluid_solid_interface_els(mod, matls)
    assert np.sum(mod["el_typ_i"] == "ASI2D2") > 0
    _cmp_assembly(mod, matls, {"dof_to_use": [1, 2, 4]})


def test_assembly_tets_matches_oracle():
    mod = PKG.mesh.cubic_structured_tet_mesh(1e-3, 0.5e-3)
    mod["el_typ_i"] = np.array(["C3D4"] * len(mod["el_mat_i"]), dtype=object)
    mod["el_abs_i"] = np.linspace(0, 1, len(mod["el_mat_i"]))
    K, C, M, gl = _cmp_assembly(mod, [steel()], {})
    # free body: rigid translations are in the null space of K
    for d in range(3):
        r = np.zeros(K.shape[0])
        r[gl[:, d] - 1] = 1.0
        assert np.abs(K @ r).max() < 1e-6 * abs(K).max()


def test_plate_workload_sizes():
    mod, matls = solid_plate(20e-3, 10e-3, 0.1e-3)
    assert mod["nds"].shape[0] == 20301
    assert mod["els"].shape[0] == 40000
    J = np.linalg.det(np.stack([mod["nds"][mod["els"][:, 1] - 1] - mod["nds"][mod["els"][:, 0] - 1],
                                mod["nds"][mod["els"][:, 2] - 1] - mod["nds"][mod["els"][:, 0] - 1]],
                               axis=1))
    assert np.all(J > 0)  # counter-clockwise, as the reference mesher emits
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2]})
    assert K.shape[0] == 40602
    assert C.nnz == 0


def test_node_numbering_matches_matlab_meshgrid():
    mod = PKG.mesh.fn_rectangular_structured_mesh(rect(2.0, 1.0), 0.5)
    # MATLAB: nodes run down y first (column-major), 3 rows x 5 columns
    assert mod["nds"].shape == (15, 2)
    np.testing.assert_allclose(mod["nds"][:4], [[0, 0], [0, 0.5], [0, 1], [0.5, 0]])
    # first 'a' element: (n(1,1), n(2,2), n(2,1)) -> 1, 5, 2 (1-based)
    np.testing.assert_array_equal(mod["els"][0], [1, 5, 2])
    np.testing.assert_array_equal(mod["els"][8], [1, 4, 5])  # first 'b' element


def test_absorbing_index_profile():
    m = PKG.mesh
    mod, _ = solid_plate(10e-3, 5e-3, 0.25e-3)
    inner = np.array([[2e-3, 2e-3], [8e-3, 2e-3], [8e-3, 50e-3], [2e-3, 50e-3]])
    mod = m.fn_add_absorbing_layer(mod, inner, 2e-3)
    c = m.fn_calc_element_centres(mod["nds"], mod["els"])
    a = mod["el_abs_i"]
    inside = (c[:, 0] > 2e-3) & (c[:, 0] < 8e-3) & (c[:, 1] > 2e-3)
    assert np.all(a[inside] == 0)
    left = (c[:, 0] < 2e-3) & (c[:, 1] > 2.5e-3)
    np.testing.assert_allclose(a[left], (2e-3 - c[left, 0]) / 2e-3, rtol=1e-12)
    assert a.max() <= 1 and a.min() >= 0


def test_interface_elements_orientation_and_count():
    mod, matls = coupled_block(2e-3, 1e-3, 0.75e-3, 0.25e-3, absorbing=False)
    asi = mod["els"][mod["el_typ_i"] == "ASI2D2"][:, :2]
    assert asi.shape[0] == 8  # 2 mm / 0.25 mm edges on the y = 1 mm interface
    nds = mod["nds"]
    a = nds[asi[:, 1] - 1] - nds[asi[:, 0] - 1]
    b = np.stack([a[:, 1], -a[:, 0]], axis=1)
    # the normal b points into the fluid (y > 1 mm)
    assert np.all(b[:, 1] > 0)
    assert np.allclose(nds[asi - 1][:, :, 1], 1e-3)


def test_driver_index_mapping_and_options_query():
    gl = np.array([[1, 2, 0, 0], [0, 0, 0, 3], [4, 5, 0, 6]])
    gi, nds, dfs, valid = PKG.fn_nds_and_dfs_to_gi([1, 2, 3, 2, 1], [2, 1, 4, 4, 3], gl)
    np.testing.assert_array_equal(gi, [2, 6, 3])
    np.testing.assert_array_equal(valid, [True, False, True, True, False])
    # gl_lookup(nds(i), dfs(i)) with an index beyond gl_lookup errors in MATLAB
    for nds_, dfs_ in (([1, 9], [1, 1]), ([1, 2], [1, 5]), ([0], [1]), ([1], [0])):
        with pytest.raises(IndexError):
            PKG.fn_nds_and_dfs_to_gi(nds_, dfs_, gl)
    opts = PKG.fn_FE_entry_point(None, None, None, {"max_damping": 1.0})
    assert opts["solver_mode"] == "vel at curent time step"
    assert opts["max_damping"] == 1.0 and opts["damping_power_law"] == 3
    assert np.isinf(opts["field_output_every_n_frames"])
    with pytest.raises(NotImplementedError):
        PKG.fn_FE_entry_point(None, None, [{}], {"solver": "pogo"})


@pytest.mark.parametrize("ftype", ["KE", "mean(u1)", "mean(v2)", "div(u)", "curl(v)"])
def test_field_output_types(ftype):
    mod, matls = coupled_block(1.5e-3, 0.75e-3, 0.5e-3, 0.25e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4]})
    rng = np.random.default_rng(2)
    f = rng.standard_normal((K.shape[0], 3))
    mats = {"M": M, "gl_lookup": gl}
    out = PKG.fn_get_field_output(f, mod, mats, ftype)
    assert out.shape == (mod["els"].shape[0], 3)
    if ftype == "KE":
        ref = op.field_output_ke(f, mod["els"], M.diagonal(), gl)
        assert rel_l2(out, ref) < 1e-14
    if ftype == "mean(u1)":
        e = 5
        nodes = [v for v in mod["els"][e] if v > 0]
        vals = [f[gl[v - 1, 0] - 1, 0] if gl[v - 1, 0] > 0 else 0.0 for v in nodes]
        assert np.isclose(out[e, 0], np.mean(vals))


def test_signals():
    t = np.linspace(0, 4e-6, 4001)
    s = PKG.mesh.hann_toneburst(t, 2e6, 5)
    assert s[t > 2.5e-6].max() == 0
    i = np.argmin(np.abs(t - 1.25e-6))
    assert np.isclose(s[i], np.sin(2 * np.pi * 2e6 * t[i]) * 1.0, atol=1e-2)
    g = PKG.mesh.fn_gaussian_pulse(t, 2e6, 5)
    assert g.shape == t.shape and abs(g).max() <= 1 and abs(g[0]) < 0.05


def test_stable_time_step_matches_dense_eigenvalue():
    mod, matls = coupled_block(1.5e-3, 0.75e-3, 0.5e-3, 0.25e-3, absorbing=False)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4]})
    m = M.diagonal()
    lam = np.max(np.real(np.linalg.eigvals(np.diag(1 / m) @ K.toarray())))
    dt = PKG.mesh.stable_time_step(K, m, safety=1.0, iters=5000, tol=1e-14)
    assert np.isclose(dt, 2 / np.sqrt(lam), rtol=1e-3)
    vmax, _ = PKG.mesh.fn_estimate_max_min_vels(matls)
    assert np.isclose(PKG.mesh.fn_get_suitable_time_step(matls, 1e-4), 1e-4 / vmax / np.sqrt(2))


@pytest.mark.parametrize("case", ["solid", "fluid_solid", "tets"])
@pytest.mark.parametrize("dofs", [[], [1, 2]])
def test_native_assembly_matches_numpy_and_oracle(case, dofs):
    m = PKG.mesh
    if case == "solid":
        mod, matls = solid_plate(2e-3, 1e-3, 0.125e-3)
        inner = np.array([[0.5e-3, 0.25e-3], [1.5e-3, 0.25e-3], [1.5e-3, 5e-3], [0.5e-3, 5e-3]])
        mod = m.fn_add_absorbing_layer(mod, inner, 0.25e-3)
    elif case == "fluid_solid":
        mod, matls = coupled_block(2e-3, 1e-3, 0.75e-3, 0.25e-3)
        keep = mod["el_typ_i"] != "ASI2D2"   # native path: no interface elements
        for k in ("els", "el_mat_i", "el_abs_i", "el_typ_i"):
            mod[k] = np.asarray(mod[k])[keep]
        if dofs:
            dofs = [1, 2, 4]
    else:
        mod = m.cubic_structured_tet_mesh(1e-3, 0.25e-3)
        mod["el_typ_i"] = np.array(["C3D4"] * len(mod["el_mat_i"]), dtype=object)
        mod["el_abs_i"] = np.linspace(0, 1, len(mod["el_mat_i"])) ** 2
        matls = [steel()]
    opts = {"dof_to_use": dofs, "damping_power_law": 2}
    K1, C1, M1, gl1 = assemble(mod, matls, dict(opts, assembly_backend="numpy"))
    K2, C2, M2, gl2 = assemble(mod, matls, dict(opts, assembly_backend="native"))
    np.testing.assert_array_equal(gl1, gl2)
    for A, B in ((K1, K2), (C1, C2), (M1, M2)):
        assert A.shape == B.shape
        assert abs(A - B).max() <= 1e-13 * max(abs(A).max(), 1e-300)
    # stored patterns may differ only where element contributions cancel to
    # rounding level in one summation order and exactly in the other
    if case != "tets" or not dofs:
        K3, C3, M3, gl3 = oa.build_global_matrices(mod["nds"], mod["els"], mod["el_mat_i"],
                                                   mod["el_abs_i"], list(mod["el_typ_i"]), matls, opts)
        np.testing.assert_array_equal(gl2, gl3)
        assert abs(K2 - K3).max() <= 1e-13 * abs(K3).max()


def test_box_absorbing_layer_3d():
    m = PKG.mesh
    mod = m.cubic_structured_tet_mesh(4e-3, 0.5e-3)
    mod = m.add_absorbing_layer_box(mod, (1e-3, 1e-3, -np.inf), (3e-3, 3e-3, np.inf), 1e-3)
    c = m.fn_calc_element_centres(mod["nds"], mod["els"])
    a = mod["el_abs_i"]
    gx = np.maximum(np.maximum(1e-3 - c[:, 0], c[:, 0] - 3e-3), 0)
    gy = np.maximum(np.maximum(1e-3 - c[:, 1], c[:, 1] - 3e-3), 0)
    np.testing.assert_allclose(a, np.clip(np.hypot(gx, gy) / 1e-3, 0, 1), rtol=0, atol=1e-15)
    inside = (gx == 0) & (gy == 0)
    assert inside.any() and np.all(a[inside] == 0) and a.max() > 0.5
    # no layer on the top and bottom faces: a does not depend on z
    top = c[:, 2] > 3.5e-3
    assert np.all(a[top & inside] == 0)


def test_tet_mesh_orientation_and_volume():
    mod = PKG.mesh.cubic_structured_tet_mesh((2e-3, 1e-3, 1.5e-3), 0.5e-3)
    X = mod["nds"][mod["els"] - 1]
    J = np.linalg.det(X[:, 1:, :] - X[:, :1, :])
    assert np.all(J > 0)
    np.testing.assert_allclose(J.sum() / 6, 2e-3 * 1e-3 * 1.5e-3, rtol=1e-12)
    assert mod["els"].shape[0] == 6 * 4 * 2 * 3


def test_interface_damping_matches_the_reference_lines():
    """fn_add_interface_damping.m:18-53: the product's diagonal form against
    the oracle's dense, line-by-line form (duplicate interface nodes included,
    as MATLAB's C(gi, gi) assignment sees them), plus what the formula fixes:
    only the interface nodes' DOFs change, each once, by level * M_ii with
    level = min sqrt(K_ii / M_ii) / pi^2 * factor; linear in the factor; a
    model without interface elements is left as it is."""
    from oracle import assembly as oa
    mod, matls = coupled_block(1.5e-3, 0.75e-3, 0.5e-3, 0.25e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    typ = np.asarray(mod["el_typ_i"], dtype=object)
    asi = np.asarray(mod["els"])[typ == "ASI2D2"]
    assert asi.shape[0] >= 2  # neighbouring interface elements share a node
    for factor in (1.0, 0.3):
        Cp = PKG.assembly.fn_add_interface_damping(mod["els"], typ, K, C, M, gl, factor)
        Co = oa.add_interface_damping(mod["els"], typ, K, C, M, gl, factor)
        np.testing.assert_allclose(Cp.toarray(), Co, rtol=1e-15, atol=0)
    dK, dM = K.diagonal(), M.diagonal()
    level = np.min(np.sqrt(dK / dM)) / np.pi ** 2
    nodes = np.unique(asi[asi > 0])
    gi = gl[nodes - 1].ravel()
    gi = np.unique(gi[gi > 0]) - 1
    D = (PKG.assembly.fn_add_interface_damping(mod["els"], typ, K, C, M, gl, 2.0) - C).toarray()
    expect = np.zeros(K.shape[0])
    expect[gi] = 2.0 * level * dM[gi]
    np.testing.assert_allclose(np.diag(D), expect, rtol=1e-14, atol=0)
    assert not (D - np.diag(np.diag(D))).any()
    # a solid-only model has no interface elements: C unchanged
    mod2, matls2 = solid_plate(1e-3, 0.5e-3, 0.25e-3)
    K2, C2, M2, gl2 = assemble(mod2, matls2, {"dof_to_use": [1, 2]})
    typ2 = np.asarray(["CPE3"] * len(mod2["els"]), dtype=object)
    C2b = PKG.assembly.fn_add_interface_damping(mod2["els"], typ2, K2, C2, M2, gl2)
    assert (C2b != C2).nnz == 0


@pytest.mark.parametrize("which", ["u", "v"])
def test_div_curl_closed_forms_on_one_element(which):
    """fn_get_field_output.m:38-58, 85-105 on one equilateral CPE3 triangle
    (centroid at c, vertices at distance R from it): a rigid translation gives
    div = curl = 0 (the unit vectors from the centroid sum to zero); the
    radial field u = x - c gives div = 3R, curl = 0; the rotation
    u = (-(y - cy), x - cx) gives div = 0, curl = -3R (the curl unit vectors
    are the outward ones turned by -90 degrees); and both are linear."""
    R = 0.4e-3
    ang = np.deg2rad([90.0, 210.0, 330.0])
    c = np.array([1e-3, 2e-3])
    nds = c + R * np.stack([np.cos(ang), np.sin(ang)], axis=1)
    mod = {"nds": nds, "els": np.array([[1, 2, 3]]), "el_mat_i": np.array([1]),
           "el_abs_i": np.zeros(1), "el_typ_i": np.array(["CPE3"], dtype=object)}
    K, C, M, gl = assemble(mod, [steel()], {"dof_to_use": [1, 2]})
    d = nds - c
    f = np.zeros((K.shape[0], 4))
    for k in range(3):
        ix, iy = gl[k, 0] - 1, gl[k, 1] - 1
        f[ix, 0], f[iy, 0] = 0.7, -0.2                 # translation
        f[ix, 1], f[iy, 1] = d[k, 0], d[k, 1]          # expansion
        f[ix, 2], f[iy, 2] = -d[k, 1], d[k, 0]         # rotation
    f[:, 3] = 2.0 * f[:, 1] - 3.0 * f[:, 2]
    mats = {"M": M, "gl_lookup": gl}
    div = PKG.fn_get_field_output(f, mod, mats, f"div({which})")
    curl = PKG.fn_get_field_output(f, mod, mats, f"curl({which})")
    # rounding: about eps times the field's size (0.7 for the translation, R else)
    atol = np.array([4e-15, 1e-15 * R, 1e-15 * R, 1e-15 * R])
    for got, want in ((div[0], [0.0, 3 * R, 0.0, 6 * R]), (curl[0], [0.0, 0.0, -3 * R, 9 * R])):
        assert np.all(np.abs(got - np.array(want)) <= atol + 1e-14 * np.abs(want)), got

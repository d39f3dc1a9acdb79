"""GPU parity of the element-by-element step (k_step_ebe) against the
oracle's literal v6 march on the same K, C, M; tolerance 1e-10 relative L2
on every monitored history and field frame (FP64), 1e-3 for FP32."""
import numpy as np
import pytest

from conftest import PKG, assemble, coupled_block, rel_l2, solid_plate
from oracle import v6 as ov6

pytestmark = pytest.mark.gpu

MODES = ["vel at curent time step", "vel at last half time step"]
TOL = 1e-10


def _mods():
    from bristolfe_v2_amd import _lib, ebe, solver
    return _lib, ebe, solver


def _case(name):
    m = PKG.mesh
    if name.startswith("coupled"):
        # solid + water with ASI2D2 interface, absorbing layer, optional void;
        # pressure source in the water and monitors in both media
        void = (m.fn_create_smooth_random_blob(0.5, 3, 48, 7) * 0.3e-3 + np.array([1.5e-3, 0.7e-3])
                if name == "coupled_void" else None)
        mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3, void=void)
        opts = {"dof_to_use": [1, 2, 4], "damping_power_law": 2}
        K, C, M, gl = assemble(mod, matls, opts)
        dt = m.stable_time_step(K, M.diagonal())
        t = np.arange(500) * dt
        F = m.hann_toneburst(t, 5e6, 3)[None, :]
        nd = m.fn_find_node_at_point(mod["nds"], [1.5e-3, 2.0e-3])
        both = np.nonzero((gl[:, 1] > 0) & (gl[:, 3] > 0))[0]
        fi = np.array([gl[nd - 1, 3] - 1, gl[both[len(both) // 3], 3] - 1])
        return mod, matls, opts, K, C, M, gl, t, fi, F
    if name == "cube":
        # at half the CFL step: at the full step dt c/m reaches 3.2 in the
        # layers, where the 'vel at last half time step' recurrence diverges
        from bristolfe_v2_amd import configs
        dt = 0.5 * configs.cube_3d(el_size=1e-3, n_t=2)["dt"]
        c = configs.cube_3d(el_size=1e-3, n_t=400, dt=dt)
        K, C, M = c["K"], c["C"], c["M"]
        return (c["mod"], c["matls"], c["fe_options"], K, C, M, c["gl_lookup"], c["time"],
                c["frc_gi"] - 1, c["frcs"])
    if name == "iso":
        from bristolfe_v2_amd import configs
        matls = [configs.steel()]
        mod = m.fn_isometric_structured_mesh(np.array([[0, 0], [4e-3, 0], [4e-3, 2e-3], [0, 2e-3]]),
                                             0.1e-3)
        mod["el_typ_i"] = np.array(["CPE3"] * len(mod["el_mat_i"]), dtype=object)
    else:
        mod, matls = solid_plate(4e-3, 2e-3, 0.1e-3)
    inner = np.array([[0.5e-3, 0.5e-3], [3.5e-3, 0.5e-3], [3.5e-3, 5e-3], [0.5e-3, 5e-3]])
    mod = m.fn_add_absorbing_layer(mod, inner, 0.5e-3)
    opts = {"dof_to_use": [] if name == "plate3" else [1, 2], "damping_power_law": 2}
    K, C, M, gl = assemble(mod, matls, opts)
    dt = m.stable_time_step(K, M.diagonal())
    t = np.arange(500) * dt
    F = m.hann_toneburst(t, 5e6, 3)[None, :]
    nd = m.fn_find_node_at_point(mod["nds"], [2e-3, 2e-3])
    fi = np.array([gl[nd - 1, 1] - 1, gl[nd - 1, 0] - 1])
    return mod, matls, opts, K, C, M, gl, t, fi, F


PATTERN = {"plate": 2, "plate3": 2, "iso": 3, "cube": 1, "coupled": 2, "coupled_void": 2}


@pytest.mark.parametrize("name", ["plate", "plate3", "iso", "cube", "coupled", "coupled_void"])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kernel", ["pattern", "generic"])
def test_ebe_parity(name, mode, kernel):
    print("case: ebe", name, mode, kernel, flush=True)
    _lib, ebe, solver = _mods()
    mod, matls, opts, K, C, M, gl, t, fi, F = _case(name)
    op = ebe.element_operator(mod, matls, opts, gl, K.shape[0])
    assert op is not None
    hi = np.unique(np.concatenate([fi, np.arange(0, K.shape[0], 97)]))
    with solver.make_plan(K, C, M, t[1] - t[0], mode, element_op=op, operator="ebe") as plan:
        info = plan.info()
        assert info["operator_kind"] == _lib.OPERATOR_EBE
        assert info["ebe_pattern"] == PATTERN[name]
        assert info["ebe_check"] < 1e-11
        assert info["ebe_groups"] == (2 if name.startswith("coupled") else 1)
        # isotropic tets on axis-aligned cells: the structurally sparse product
        assert info["ebe_sparse"] == (1 if name == "cube" else 0)
        if name.startswith("coupled"):
            assert info["n_interface"] > 0
        plan.set_tuning(_lib.TUNE_EBE_KERNEL, 1 if kernel == "pattern" else 0)
        h, fld, ti = plan.run(t.size, fi, F, hi, 50, 1)
    r, rf, _, rti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi + 1, F, hi + 1, 50, "KE", mode)
    assert ti == rti
    assert np.linalg.norm(r) > 0
    assert rel_l2(h, r) < TOL
    assert rel_l2(fld, rf) < TOL


def test_ebe_and_csr_runs_agree_and_switch():
    print("case: ebe vs csr switch", flush=True)
    _lib, ebe, solver = _mods()
    mod, matls, opts, K, C, M, gl, t, fi, F = _case("cube")
    op = ebe.element_operator(mod, matls, opts, gl, K.shape[0])
    hi = np.arange(0, K.shape[0], 53)
    with solver.make_plan(K, C, M, t[1] - t[0], element_op=op, operator="ebe") as plan:
        h_e, _, _ = plan.run(t.size, fi, F, hi)
        plan.set_tuning(_lib.TUNE_OPERATOR, _lib.OPERATOR_CSR)
        assert plan.info()["operator_kind"] == _lib.OPERATOR_CSR
        h_c, _, _ = plan.run(t.size, fi, F, hi)
        plan.set_tuning(_lib.TUNE_OPERATOR, _lib.OPERATOR_EBE)
        h_e2, _, _ = plan.run(t.size, fi, F, hi)
    assert rel_l2(h_e, h_c) < 1e-12
    np.testing.assert_array_equal(h_e, h_e2)  # deterministic


def test_wrong_templates_are_rejected():
    print("case: ebe rejection", flush=True)
    _lib, ebe, solver = _mods()
    mod, matls, opts, K, C, M, gl, t, fi, F = _case("plate")
    op = ebe.element_operator(mod, matls, opts, gl, K.shape[0])
    op.tmpl[1] *= 1.0 + 1e-6
    with pytest.raises(_lib.BfeError) as ei:
        solver.make_plan(K, C, M, t[1] - t[0], element_op=op, operator="ebe")
    assert ei.value.status == _lib.BFE_ERR_ARG
    # "auto" falls back to CSR when the library rejects the element operator
    # (the model is below the automatic size threshold, so lower it here)
    threshold = solver.AUTO_EBE_MIN_DOF
    solver.AUTO_EBE_MIN_DOF = 0
    try:
        with solver.make_plan(K, C, M, t[1] - t[0], element_op=op, operator="auto") as plan:
            assert plan.info()["operator_kind"] == _lib.OPERATOR_CSR
    finally:
        solver.AUTO_EBE_MIN_DOF = threshold


def test_ebe_single_precision_error_reported():
    print("case: ebe fp32", flush=True)
    _lib, ebe, solver = _mods()
    mod, matls, opts, K, C, M, gl, t, fi, F = _case("cube")
    op = ebe.element_operator(mod, matls, opts, gl, K.shape[0])
    hi = np.arange(0, K.shape[0], 101)
    with solver.make_plan(K, C, M, t[1] - t[0], solver_precision="single", element_op=op,
                          operator="ebe") as plan:
        h, _, _ = plan.run(t.size, fi, F, hi)
    r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi + 1, F, hi + 1)
    err = rel_l2(h, r)
    print(f"  fp32 ebe rel L2 vs fp64 oracle: {err:.2e}")
    assert err < 1e-3

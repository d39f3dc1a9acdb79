"""GPU tests of the batched (multi-source) march beyond the single-model
FP64 cases: the FP32 bulk march against the FP64 oracle, several models of
uneven size marched as one block-diagonal system with 64 source columns
(every column against its own single-source oracle run), and forcing with one
signal row per forcing entry. The reference marches each source on its own
(subdoms/fn_run_main_model.m:101-115 calling fn_BristolFE_v2.m:72-123)."""
import numpy as np
import pytest

from conftest import PKG, assemble, coupled_block, rel_l2
from oracle import v6 as ov6

pytestmark = pytest.mark.gpu

MODES = ["vel at curent time step", "vel at last half time step"]
TOL = 1e-10
TOL_FP32 = 1e-3   # FP32 bulk march (FP64 set-up) against the FP64 oracle


def _coupled(n_t=300):
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    t = np.arange(n_t) * dt
    return mod, K, C, M, gl, dt, t


@pytest.mark.parametrize("mode", MODES)
def test_batched_single_precision_error_bounded(mode):
    """run_batched with solver_precision 'single' on a damped coupled model
    (interface rows, absorbing layer), 16 sources: every column against the
    FP64 oracle below 1e-3 relative L2; the error is printed."""
    print("case: batched fp32", mode, flush=True)
    mod, K, C, M, gl, dt, t = _coupled()
    s = PKG.mesh.hann_toneburst(t, 5e6, 3)
    src = gl[np.nonzero(gl[:, 3] > 0)[0][::7][:16], 3] - 1
    S = src.size
    hi = np.concatenate([src, np.arange(0, K.shape[0], 23)])
    plan = PKG.make_plan(K, C, M, dt, mode, "single")
    try:
        assert plan.info()["precision"] == 1
        h32, _ = plan.run_batched(t.size, S, src, np.arange(S), s, hi)
    finally:
        plan.close()
    A, B, D = ov6.abd_matrices(K, C, M, dt, mode)
    worst = 0.0
    for k in (0, S // 2, S - 1):
        r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [src[k] + 1], s, hi + 1,
                                                    solver_mode=mode, abd=(A, B, D))
        assert np.linalg.norm(r) > 0
        worst = max(worst, rel_l2(h32[:, k, :], r))
    print(f"batched fp32 worst column rel L2 vs fp64 oracle: {worst:.3e}")
    assert worst < TOL_FP32


def test_multimodel_block_diagonal_launch():
    """A reduced form of the FMC workload: a 64-element array (0.1 mm pitch)
    on an 8 mm x 4 mm plate plus three seeded subdomain models of uneven size
    (absorbing rims), one block-diagonal system marched with 64 columns in one
    launch sequence. Every column of every model matches the oracle's
    single-source run of the same forcing entries to 1e-10."""
    print("case: multi-model block-diagonal launch", flush=True)
    from bristolfe_v2_amd import configs
    c = configs.fmc_multimodel(pitch=0.1e-3, n_sub=3, sub_dof=(4e3, 12e3), n_t=250,
                               plate=dict(width=8e-3, height=4e-3), sub_height=4e-3)
    K, C, M, t = c["K"], c["C"], c["M"], c["time"]
    sizes = c["model_sizes"]
    assert len(sizes) == 4 and len(set(sizes)) == 4
    S = c["n_src"]
    fd, fs, hd, F = c["frc_dof"], c["frc_src"], c["hist_dof"], c["frcs"]
    plan = PKG.make_plan(K, C, M, c["dt"])
    try:
        hb, ti = plan.run_batched(t.size, S, fd, fs, F, hd)
        assert plan.info()["damped"] == 1
    finally:
        plan.close()
    assert ti == 1   # one broadcast signal row (v6:163-166)
    A, B, D = ov6.abd_matrices(K, C, M, c["dt"], "vel at curent time step")
    worst = 0.0
    for k in range(S):
        sel = fs == k
        r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fd[sel] + 1, F, hd + 1,
                                                    abd=(A, B, D))
        # every model is excited in every column
        offs = c["offsets"]
        for b in range(len(sizes)):
            rows = (hd >= offs[b]) & (hd < offs[b + 1])
            assert np.linalg.norm(r[rows]) > 0
        worst = max(worst, rel_l2(hb[:, k, :], r))
    print(f"multi-model worst column rel L2 vs oracle: {worst:.3e}")
    assert worst < TOL


@pytest.mark.parametrize("mode", MODES)
def test_batched_per_entry_signal_rows(mode):
    """Forcing with one signal row per forcing entry (p rows, each source
    column with three entries of different delays and amplitudes, two
    columns sharing a DOF): the batched march matches the single-source
    runs and the oracle, and column tiles (autotune.run_tiled with a tile
    narrower than the sources) give the same result."""
    print("case: batched per-entry signal rows", mode, flush=True)
    from bristolfe_v2_amd import autotune
    mod, K, C, M, gl, dt, t = _coupled()
    water = np.nonzero(gl[:, 3] > 0)[0]
    S = 10
    fd, fs, rows = [], [], []
    for k in range(S):
        for j in range(3):
            nd = water[(37 * k + 11 * j) % water.size]
            fd.append(gl[nd, 3] - 1)
            fs.append(k)
            rows.append((1 + 0.3 * j) * PKG.mesh.hann_toneburst(t - (5 * k + 9 * j) * dt, 5e6, 3))
    fd[3] = fd[0]   # columns 0 and 1 share a forcing DOF
    fd, fs, F = np.array(fd), np.array(fs), np.vstack(rows)
    hi = np.concatenate([fd, np.arange(0, K.shape[0], 19)])
    plan = PKG.make_plan(K, C, M, dt, mode)
    try:
        hb, ti = plan.run_batched(t.size, S, fd, fs, F, hi)
        # p rows: the march starts at the first column of F with a non-zero entry
        assert ti == ov6.ti_start_of(F) > 1
        for k in range(S):
            sel = fs == k
            h1, _, _ = plan.run(t.size, fd[sel], F[sel], hi)
            assert rel_l2(hb[:, k, :], h1) < 1e-13
        ht, _ = autotune.run_tiled(plan, t.size, S, fd, fs, F, hi, 4)
        assert rel_l2(ht, hb) < 1e-13
    finally:
        plan.close()
    A, B, D = ov6.abd_matrices(K, C, M, dt, mode)
    for k in (0, S - 1):
        sel = fs == k
        r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fd[sel] + 1, F[sel], hi + 1,
                                                    solver_mode=mode, abd=(A, B, D))
        assert np.linalg.norm(r) > 0
        assert rel_l2(hb[:, k, :], r) < TOL


def _ebe_case(name):
    """(K, C, M, dt, t, op, sources, monitors) of a small model whose element
    operator has a mesh-pattern kernel."""
    from bristolfe_v2_amd import configs, ebe
    m = PKG.mesh
    if name == "coupled":
        # two groups (CPE3 solid, AC2D3 water; a node -> first-DOF table),
        # ASI2D2 interface rows, absorbing layer
        mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3)
        opts = {"dof_to_use": [1, 2, 4], "damping_power_law": 2}
        K, C, M, gl = assemble(mod, matls, opts)
        src = gl[np.nonzero(gl[:, 3] > 0)[0][::5][:24], 3] - 1
    elif name == "cube":
        c = configs.cube_3d(el_size=1e-3, n_t=10)
        mod, matls, opts = c["mod"], c["matls"], c["fe_options"]
        K, C, M, gl = c["K"], c["C"], c["M"], c["gl_lookup"]
        top = np.nonzero(np.isclose(mod["nds"][:, 2], mod["nds"][:, 2].max()))[0]
        src = gl[top[::3][:20], 2] - 1
    else:  # the union of disjoint plate meshes of the FMC workload, reduced
        c = configs.fmc_multimodel(pitch=0.1e-3, n_sub=3, sub_dof=(4e3, 12e3), n_t=10,
                                   plate=dict(width=8e-3, height=4e-3), sub_height=4e-3)
        mod, matls, opts = c["mod"], c["matls"], c["fe_options"]
        K, C, M, gl = c["K"], c["C"], c["M"], c["gl_lookup"]
        src = c["frc_dof"][c["frc_src"] < 64][::4]
    op = ebe.element_operator(mod, matls, opts, gl, K.shape[0])
    assert op is not None
    dt = m.stable_time_step(K, M.diagonal())
    if name == "cube":  # dt c/m < 2 in the layers: both solver modes stable
        dt *= 0.5
    return K, C, M, dt, op, np.asarray(src), np.arange(0, K.shape[0], 13)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", ["union", "coupled", "cube"])
def test_batched_element_product(name, mode):
    """The multi-source element product (k_step_ebe_multi) against the CSR
    multi-source step on the same plan (1e-13) and every 4th column against
    the oracle's single-source run (1e-10): the union of disjoint plate meshes
    (irregular nodes at the mesh seams), the coupled solid/water model (two
    groups, a node -> DOF table, interface rows) and a C3D4 cube (stencil and
    structurally sparse slot paths). Nodes per wavefront from 1 to 64, with
    and without the sliding neighbour rows, give the same bits. The node-block
    form of the triangle groups matches CSR to 1e-13 as well."""
    print("case: batched element product", name, mode, flush=True)
    from bristolfe_v2_amd import _lib
    K, C, M, dt, op, src, hi = _ebe_case(name)
    n_t = 200 if name != "cube" else 120
    t = np.arange(n_t) * dt
    s = PKG.mesh.hann_toneburst(t, 5e6 if name != "cube" else 2e6, 3)
    S = src.size
    plan = PKG.make_plan(K, C, M, dt, mode, element_op=op, operator="ebe")
    try:
        assert plan.info()["operator_kind"] == 1
        hc, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
        assert plan.info()["multi_operator"] == 0   # CSR by default
        plan.set_tuning(_lib.TUNE_MULTI_OPERATOR, 1)
        he, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
        assert plan.info()["multi_operator"] == 1
        assert np.linalg.norm(hc) > 0
        assert rel_l2(he, hc) < 1e-13
        for npw, slide in ((1, 1), (2, 1), (3, 1), (64, 1), (4, 0)):
            plan.set_tuning(_lib.TUNE_MULTI_NODES_PER_WAVE, npw)
            plan.set_tuning(_lib.TUNE_MULTI_SLIDE, slide)
            h2, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
            np.testing.assert_array_equal(h2, he)
        # the node-block form (triangle patterns; the tets fall back to CSR)
        plan.set_tuning(_lib.TUNE_MULTI_OPERATOR, _lib.OPERATOR_NODE)
        hn, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
        if name == "cube":
            assert plan.info()["multi_operator"] == _lib.OPERATOR_CSR
            np.testing.assert_array_equal(hn, hc)
        else:
            assert plan.info()["multi_operator"] == _lib.OPERATOR_NODE
            assert rel_l2(hn, hc) < 1e-13
            for npw in (1, 3, 64):
                plan.set_tuning(_lib.TUNE_MULTI_NODES_PER_WAVE, npw)
                h2, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
                np.testing.assert_array_equal(h2, hn)
    finally:
        plan.close()
    A, B, D = ov6.abd_matrices(K, C, M, dt, mode)
    for k in range(0, S, 4):
        r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [src[k] + 1], s, hi + 1,
                                                    solver_mode=mode, abd=(A, B, D))
        assert rel_l2(he[:, k, :], r) < TOL


def test_batched_element_product_single_precision():
    """FP32 bulk march on the multi-source element product: error against the
    FP64 oracle below 1e-3 (union of plate meshes, 16 sources)."""
    print("case: batched element product fp32", flush=True)
    from bristolfe_v2_amd import _lib
    K, C, M, dt, op, src, hi = _ebe_case("union")
    src = src[:16]
    n_t = 200
    t = np.arange(n_t) * dt
    s = PKG.mesh.hann_toneburst(t, 5e6, 3)
    plan = PKG.make_plan(K, C, M, dt, "vel at curent time step", "single", element_op=op,
                         operator="ebe")
    try:
        plan.set_tuning(_lib.TUNE_MULTI_OPERATOR, 1)
        h32, _ = plan.run_batched(n_t, src.size, src, np.arange(src.size), s, hi)
        assert plan.info()["multi_operator"] == 1
    finally:
        plan.close()
    A, B, D = ov6.abd_matrices(K, C, M, dt, "vel at curent time step")
    worst = 0.0
    for k in (0, src.size - 1):
        r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [src[k] + 1], s, hi + 1,
                                                    abd=(A, B, D))
        worst = max(worst, rel_l2(h32[:, k, :], r))
    print(f"batched element product fp32 rel L2 vs fp64 oracle: {worst:.3e}")
    assert worst < TOL_FP32

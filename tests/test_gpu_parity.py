"""GPU parity: the HIP march (through the C ABI) against the oracle's literal
restatement of fn_explicit_dynamic_solver_v6.m on the same inputs.

Tolerance: 1e-10 relative L2 on every monitored history (FP64 path); the
FP32 path reports its error against the FP64 oracle and must stay below 1e-3.
"""
import numpy as np
import pytest

from conftest import PKG, assemble, coupled_block, rel_l2, solid_plate
from oracle import v6 as ov6

pytestmark = pytest.mark.gpu

MODES = ["vel at curent time step", "vel at last half time step"]
TOL = 1e-10


def _setup(mod, matls, opts, n_t=400, f0=5e6, src=None, dof=2):
    K, C, M, gl = assemble(mod, matls, opts)
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    t = np.arange(n_t) * dt
    F = PKG.mesh.hann_toneburst(t, f0, 3)
    nds = mod["nds"]
    if src is None:
        src = nds.mean(axis=0)
    nd = PKG.mesh.fn_find_node_at_point(nds, src)
    fi = np.array([gl[nd - 1, dof - 1]])
    assert fi[0] > 0
    hi = np.arange(1, K.shape[0] + 1, max(1, K.shape[0] // 97))
    return K, C, M, gl, t, F, fi, hi


def _gpu(K, C, M, t, fi, F, hi, every=np.inf, ftype="KE", mode=MODES[0], prec="double"):
    return PKG.fn_explicit_dynamic_solver_v6(K, C, M, t, fi, F, [], [], hi, every, 1, ftype, mode, prec)


@pytest.mark.parametrize("mode", MODES)
def test_solid_plate_parity(mode):
    print("case: solid plate", mode, flush=True)
    mod, matls = solid_plate(4e-3, 2e-3, 0.1e-3)
    K, C, M, gl, t, F, fi, hi = _setup(mod, matls, {"dof_to_use": [1, 2]}, 500)
    h, fld, _, ft = _gpu(K, C, M, t, fi, F, hi, 50, "KE", mode)
    r, rf, rft, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, 50, "KE", mode)
    assert rel_l2(h, r) < TOL
    assert rel_l2(fld, rf) < TOL
    np.testing.assert_array_equal(ft, rft)


@pytest.mark.parametrize("mode", MODES)
def test_coupled_absorbing_parity(mode):
    print("case: coupled + absorbing", mode, flush=True)
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3)
    K, C, M, gl, t, F, fi, hi = _setup(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2},
                                      600, 5e6, src=[1.5e-3, 2.0e-3], dof=4)
    plan = PKG.make_plan(K, C, M, t[1] - t[0], mode)
    info = plan.info()
    plan.close()
    assert info["n_interface"] > 0 and info["damped"] == 1
    h, fld, _, ft = _gpu(K, C, M, t, fi, F, hi, 100, "mean(u1)", mode)
    r, rf, rft, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, 100, "mean(u1)", mode)
    assert np.linalg.norm(r) > 0
    assert rel_l2(h, r) < TOL
    assert rel_l2(fld, rf) < TOL


def test_coupled_void_all_dofs():
    print("case: coupled void, dof_to_use all", flush=True)
    void = PKG.mesh.fn_create_smooth_random_blob(0.5, 3, 48, 7) * 0.3e-3 + np.array([1.5e-3, 0.7e-3])
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3, void=void)
    K, C, M, gl, t, F, fi, hi = _setup(mod, matls, {}, 500, 5e6, src=[1.5e-3, 2.0e-3], dof=4)
    h, _, _, _ = _gpu(K, C, M, t, fi, F, hi)
    r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi)
    assert rel_l2(h, r) < TOL


def test_forcing_semantics_on_gpu():
    """Broadcast row, per-row signals with a duplicate DOF, leading zeros
    (ti_start) and an interface forcing DOF."""
    print("case: forcing semantics", flush=True)
    mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4]})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n_t = 300
    t = np.arange(n_t) * dt
    s = PKG.mesh.hann_toneburst(t - 25 * dt, 5e6, 3)
    # an interface node (carries u1, u2 and p)
    both = np.nonzero((gl[:, 1] > 0) & (gl[:, 3] > 0))[0]
    nd_i = both[len(both) // 2] + 1
    fi = np.array([gl[nd_i - 1, 3], gl[3, 1], gl[nd_i - 1, 3], gl[5, 0]])
    assert np.all(fi > 0)
    F = np.vstack([s, 0.5 * s, -2 * s, np.roll(s, 9)])
    hi = np.concatenate([fi, np.arange(1, K.shape[0] + 1, 13)])
    for rows in (F, s[None, :]):
        h, _, _, _ = _gpu(K, C, M, t, fi, rows, hi)
        r, _, _, ti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, rows, hi)
        # one row: v6's sum(abs(F)) is a scalar, so ti_start = 1
        assert ti > 20 if rows.shape[0] > 1 else ti == 1
        assert not h[:, :25].any()
        assert rel_l2(h, r) < TOL


def test_empty_and_minimal_marches():
    """Edge cases of the march against the oracle: no forcing entry (an
    empty loop: every output 0), no monitor (field frames only), the
    shortest marches (2 and 3 samples; v6 needs two for dt), a one-row
    signal non-zero only at its last sample, and a p-row forcing that is all
    zero (an empty loop again)."""
    print("case: empty and minimal marches", flush=True)
    from bristolfe_v2_amd import solver
    mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n = K.shape[0]
    fi = np.array([gl[3, 1] - 1, gl[7, 0] - 1])
    hi = np.arange(0, n, 11)
    with solver.make_plan(K, C, M, dt) as plan:
        for n_t in (2, 3, 60):
            t = np.arange(n_t) * dt
            s = PKG.mesh.hann_toneburst(t, 5e6, 3)
            cases = [
                ("no forcing", np.zeros(0, np.int64), np.zeros((1, n_t)), hi),
                ("no monitors", fi, s[None, :], np.zeros(0, np.int64)),
                ("last sample only", fi[:1], np.eye(1, n_t, n_t - 1), hi),
                ("all-zero rows", fi, np.zeros((2, n_t)), hi),
                ("plain", fi, s[None, :], hi),
            ]
            for name, f_i, F, h_i in cases:
                h, fld, ti = plan.run(n_t, f_i, F, h_i, 1, 1)
                r, rf, _, rti = ov6.explicit_dynamic_solver_v6(K, C, M, t, f_i + 1, F, h_i + 1, 1)
                assert h.shape == (h_i.size, n_t) and fld.shape == (n, n_t), name
                assert ti == (rti if rti is not None else n_t + 1), (name, ti, rti)
                if not np.any(r) and not np.any(rf):
                    assert not h.any() and not fld.any(), name
                else:
                    assert rel_l2(h, r) < TOL and rel_l2(fld, rf) < TOL, name


def test_batched_edge_cases():
    """Batched marches with a column that no entry forces (all zero), with
    no monitor, with 64 columns (the most one launch takes) and delays that
    differ per column (each column's own first forced step), and the
    library's refusal of 65 columns."""
    print("case: batched edge cases", flush=True)
    from bristolfe_v2_amd import _lib, solver
    mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n, n_t = K.shape[0], 120
    t = np.arange(n_t) * dt
    rng = np.random.default_rng(9)
    hi = np.arange(0, n, 9)
    with solver.make_plan(K, C, M, dt) as plan:
        # three columns, the middle one unforced
        fd = np.array([5, 40, 77])
        fs = np.array([0, 2, 2])
        F = np.vstack([PKG.mesh.hann_toneburst(t - k * dt, 5e6, 3) for k in (0, 10, 30)])
        hb, _ = plan.run_batched(n_t, 3, fd, fs, F, hi)
        assert not hb[:, 1, :].any() and hb[:, 0, :].any() and hb[:, 2, :].any()
        h0, _, _ = plan.run(n_t, fd[:1], F[:1], hi)
        assert rel_l2(hb[:, 0, :], h0) < 1e-13
        h2, _, _ = plan.run(n_t, fd[1:], F[1:], hi)
        assert rel_l2(hb[:, 2, :], h2) < 1e-13
        # no monitor
        hq, _ = plan.run_batched(n_t, 3, fd, fs, F, np.zeros(0, np.int64))
        assert hq.shape == (0, 3, n_t)
        # 64 columns, one entry each with its own delay
        S = 64
        fd64 = rng.choice(n, S, replace=False)
        F64 = np.vstack([PKG.mesh.hann_toneburst(t - int(rng.integers(0, 40)) * dt, 5e6, 3)
                         for _ in range(S)])
        hb64, _ = plan.run_batched(n_t, S, fd64, np.arange(S), F64, hi)
        for s in (0, 17, 63):
            r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fd64[s:s + 1] + 1,
                                                        F64[s:s + 1], hi + 1)
            assert rel_l2(hb64[:, s, :], r) < TOL
        with pytest.raises(_lib.BfeError):
            plan.run_batched(n_t, 65, fd64[:1], np.zeros(1, np.int64), F64[:1], hi)


@pytest.mark.parametrize("mode", MODES)
def test_delayed_single_row_field_times(mode):
    """A broadcast toneburst delayed by 40 steps with velocity frames every 7
    steps: v6:163-166 marches a single-row forcing from ti = 1, so every frame,
    including those before the signal, carries time(ti) (v6:200-208); the
    field times must equal the oracle's exactly, the frames and histories to
    1e-10."""
    print("case: delayed single-row field times", mode, flush=True)
    mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n_t = 240
    t = np.arange(n_t) * dt
    s = PKG.mesh.hann_toneburst(t - 40 * dt, 5e6, 3)
    assert not s[:41].any()
    nds = np.nonzero(gl[:, 3] > 0)[0][::37] + 1
    fi = gl[nds - 1, 3]
    hi = np.concatenate([fi, np.arange(1, K.shape[0] + 1, 11)])
    h, fld, _, ft = _gpu(K, C, M, t, fi, s, hi, 7, "KE", mode)
    r, rf, rft, ti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, s, hi, 7, "KE", mode)
    assert ti == 1
    frames = np.arange(1, n_t + 1, 7)
    np.testing.assert_array_equal(rft, t[frames - 1])
    np.testing.assert_array_equal(ft, rft)
    assert rel_l2(h, r) < TOL
    assert rel_l2(fld, rf) < TOL
    plan = PKG.make_plan(K, C, M, dt, mode)
    plan.run(n_t, fi - 1, s[None, :], hi - 1)
    assert plan.stats()["steps"] == n_t
    plan.close()


def test_zero_forcing_outputs_zero():
    mod, matls = solid_plate(1e-3, 1e-3, 0.25e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2]})
    t = np.arange(50) * 1e-9
    h, fld, _, ft = _gpu(K, C, M, t, [3], np.zeros(50), [3, 4], 10)
    assert not h.any() and not fld.any() and not ft.any()


def test_single_precision_error_reported():
    print("case: fp32", flush=True)
    mod, matls = solid_plate(4e-3, 2e-3, 0.1e-3)
    K, C, M, gl, t, F, fi, hi = _setup(mod, matls, {"dof_to_use": [1, 2]}, 500)
    h32, _, _, _ = _gpu(K, C, M, t, fi, F, hi, prec="single")
    r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi)
    err = rel_l2(h32, r)
    print(f"fp32 relative L2 vs fp64 oracle: {err:.3e}")
    assert err < 1e-3


def test_driver_end_to_end():
    """fn_FE_entry_point with the reference's struct fields."""
    print("case: driver", flush=True)
    mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4]})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    t = np.arange(300) * dt
    nd = PKG.mesh.fn_find_node_at_point(mod["nds"], [1e-3, 1.5e-3])
    steps = {"load": {"time": t, "frcs": PKG.mesh.hann_toneburst(t, 5e6, 3), "frc_nds": [nd],
                      "frc_dfs": [4]},
             "mon": {"nds": [nd, nd, 1, 2], "dfs": [4, 1, 1, 2]}}
    opts = {"dof_to_use": [1, 2, 4], "field_output_every_n_frames": 50}
    res, mats = PKG.fn_FE_entry_point(mod, matls, steps, opts, return_mats=True)
    r = res[0]
    assert list(r["valid_mon_dsps"]) == [True, False, True, True]  # water node has no u1
    gi = np.array([gl[nd - 1, 3], gl[0, 0], gl[1, 1]])
    ref, rf, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [gl[nd - 1, 3]], steps["load"]["frcs"],
                                                   gi, 50, "KE")
    assert rel_l2(r["dsps"][[0, 2, 3]], ref) < TOL
    assert not r["dsps"][1].any()
    from oracle import post as op
    ke = op.field_output_ke(rf, mod["els"], M.diagonal(), gl)
    assert r["fld"].shape == ke.shape
    assert rel_l2(r["fld"], ke) < TOL


@pytest.mark.parametrize("mode", MODES)
def test_batched_matches_single_runs(mode):
    print("case: batched", mode, flush=True)
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n_t = 300
    t = np.arange(n_t) * dt
    s = PKG.mesh.hann_toneburst(t, 5e6, 3)
    top = np.nonzero(gl[:, 3] > 0)[0][:12] + 1
    src = gl[top - 1, 3]
    hi = np.concatenate([src, np.arange(1, K.shape[0] + 1, 31)])
    plan = PKG.make_plan(K, C, M, dt, mode)
    try:
        hb, ti = plan.run_batched(n_t, src.size, src - 1, np.arange(src.size), s, hi - 1)
        for k in range(src.size):
            h1, _, _ = plan.run(n_t, [src[k] - 1], s, hi - 1)
            assert rel_l2(hb[:, k, :], h1) < 1e-13
    finally:
        plan.close()
    # reciprocity of the full-matrix capture
    fmc = hb[:src.size]  # (rx, tx, t)
    assert rel_l2(fmc, fmc.transpose(1, 0, 2)) < 1e-10


def test_plate_workload_full_size_parity():
    """The 20 mm x 10 mm, 0.1 mm plate (40,602 DOF), 3000 steps, field every 100."""
    print("case: plate_2d full size", flush=True)
    from bristolfe_v2_amd import configs
    c = configs.plate_2d()
    K, C, M, t = c["K"], c["C"], c["M"], c["time"]
    assert K.shape[0] == 40602
    h, fld, _, _ = _gpu(K, C, M, t, c["frc_gi"], c["frcs"], c["hist_gi"], 100)
    r, rf, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, c["frc_gi"], c["frcs"], c["hist_gi"], 100)
    err = rel_l2(h, r)
    print(f"plate_2d rel L2 = {err:.3e}")
    assert err < TOL
    assert rel_l2(fld, rf) < TOL


def test_batched_tuning_does_not_change_results():
    """Every launch shape of the multi-source kernel gives bit-identical
    histories (wavefronts per workgroup, rows per wavefront, dispatch order),
    and column tiles agree."""
    print("case: batched tuning", flush=True)
    from bristolfe_v2_amd import _lib, autotune
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n_t = 200
    s = PKG.mesh.hann_toneburst(np.arange(n_t) * dt, 5e6, 3)
    src = gl[np.nonzero(gl[:, 3] > 0)[0][:40], 3] - 1
    S = src.size
    hi = np.concatenate([src, np.arange(0, K.shape[0], 17)])
    plan = PKG.make_plan(K, C, M, dt)
    try:
        ref, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
        for rows in (1, 2, 8):
            for remap in (0, 1):
                for rpw in (1, 2, 8):
                    plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_BLOCK, rows)
                    plan.set_tuning(_lib.TUNE_XCD_REMAP, remap)
                    plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_WAVE, rpw)
                    h, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
                    np.testing.assert_array_equal(h, ref)
        h, _ = autotune.run_tiled(plan, n_t, S, src, np.arange(S), s, hi, 16)
        assert rel_l2(h, ref) < 1e-13
    finally:
        plan.close()


@pytest.mark.parametrize("absorbing", [False, True])
def test_batched_row_groups_agree(absorbing):
    """2D solid plate (two rows per node whose column lists differ only by
    exact zeros): the row-grouped gathers of the multi-source kernel (groups
    of 2, 4 and 8 rows over the union of their columns) agree with one gather
    per row to rounding, for row counts per wavefront from the group size to
    16 (and a ragged last slice), and the batched columns match the oracle.
    The automatic choice takes groups of 4 rows (two nodes), at most the
    rows per wavefront."""
    print("case: batched row groups, absorbing", absorbing, flush=True)
    from bristolfe_v2_amd import _lib
    mod, matls = solid_plate(1.6e-3, 0.9e-3, 0.1e-3)
    if absorbing:
        inner = np.array([[0.3e-3, 0.3e-3], [1.3e-3, 0.3e-3], [1.3e-3, 5e-3], [0.3e-3, 5e-3]])
        mod = PKG.mesh.fn_add_absorbing_layer(mod, inner, 0.3e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2], "damping_power_law": 2})
    n = K.shape[0]
    assert n % 64 != 0  # the last SELL slice is ragged
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n_t = 150
    s = PKG.mesh.hann_toneburst(np.arange(n_t) * dt, 5e6, 3)
    top = np.nonzero(np.isclose(mod["nds"][:, 1], 0.9e-3))[0]
    src = gl[top[:24], 1] - 1
    S = src.size
    hi = np.concatenate([src, np.arange(0, n, 13)])
    plan = PKG.make_plan(K, C, M, dt)
    try:
        plan.set_tuning(_lib.TUNE_MULTI_GROUP, 1)
        ref, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
        assert plan.info()["multi_group"] == 1
        for g in (0, 2, 4, 8):
            plan.set_tuning(_lib.TUNE_MULTI_GROUP, g)
            for rpw in sorted({max(g, 2), 8, 16}):
                for rows in (1, 4):
                    plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_WAVE, rpw)
                    plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_BLOCK, rows)
                    h, _ = plan.run_batched(n_t, S, src, np.arange(S), s, hi)
                    # automatic: groups of 4 (two nodes) on a 2D solid
                    assert plan.info()["multi_group"] == (g or min(4, rpw))
                    assert rel_l2(h, ref) < 1e-13
        for c in (0, S - 1):
            r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, np.arange(n_t) * dt,
                                                        np.array([src[c] + 1]), s, hi + 1)
            assert np.linalg.norm(r[c]) > 0
            assert rel_l2(h[:, c, :], r) < TOL
    finally:
        plan.close()


@pytest.mark.parametrize("mode", MODES)
def test_row_groups_automatic_choice(mode):
    """Automatic grouping pairs rows only when that saves at least a third of
    the gathers: the coupled solid/water model (nodes of 1, 2 and 3 DOFs, so
    the pairs straddle nodes) keeps one gather per row, and an explicit group
    is still honoured there (and agrees, interface rows included)."""
    from bristolfe_v2_amd import _lib
    mod, matls = coupled_block(1.5e-3, 0.75e-3, 0.5e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    s = PKG.mesh.hann_toneburst(np.arange(100) * dt, 5e6, 3)
    src = gl[np.nonzero(gl[:, 3] > 0)[0][:8], 3] - 1
    hi = np.arange(0, K.shape[0], 7)
    plan = PKG.make_plan(K, C, M, dt, mode)
    try:
        plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_WAVE, 8)
        ref, _ = plan.run_batched(100, src.size, src, np.arange(src.size), s, hi)
        assert plan.info()["multi_group"] == 1
        plan.set_tuning(_lib.TUNE_MULTI_GROUP, 4)
        h, _ = plan.run_batched(100, src.size, src, np.arange(src.size), s, hi)
        assert plan.info()["multi_group"] == 4
        assert rel_l2(h, ref) < 1e-13
    finally:
        plan.close()


@pytest.mark.parametrize("model", ["coupled", "plate"])
@pytest.mark.parametrize("mode", MODES)
def test_persistent_march_matches_eager(model, mode):
    """The persistent multi-step kernel (one cooperative launch for the whole
    loop, grid barriers between steps and after the interface solve) gives
    the eager march's bits: the coupled model (interface rows, absorbing
    layer) and an absorbing plate, forcing with leading zeros, monitors and
    velocity field frames; and the oracle's histories."""
    print("case: persistent march", model, mode, flush=True)
    from bristolfe_v2_amd import _lib
    if model == "coupled":
        mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
        K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
        nd = PKG.mesh.fn_find_node_at_point(mod["nds"], [1e-3, 1.5e-3])
        fi = np.array([gl[nd - 1, 3] - 1, gl[3, 1] - 1])
    else:
        mod, matls = solid_plate(2e-3, 1e-3, 0.1e-3)
        inner = np.array([[0.3e-3, 0.3e-3], [1.7e-3, 0.3e-3], [1.7e-3, 5e-3], [0.3e-3, 5e-3]])
        mod = PKG.mesh.fn_add_absorbing_layer(mod, inner, 0.3e-3)
        K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2], "damping_power_law": 2})
        nd = PKG.mesh.fn_find_node_at_point(mod["nds"], [1e-3, 1e-3])
        fi = np.array([gl[nd - 1, 1] - 1, gl[nd - 1, 0] - 1])
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n_t = 300
    t = np.arange(n_t) * dt
    F = np.vstack([PKG.mesh.hann_toneburst(t - 25 * dt, 5e6, 3),
                   0.5 * PKG.mesh.hann_toneburst(t - 40 * dt, 4e6, 2)])
    hi = np.unique(np.concatenate([fi, np.arange(0, K.shape[0], 29)]))
    plan = PKG.make_plan(K, C, M, dt, mode)
    try:
        assert (plan.info()["n_interface"] > 0) == (model == "coupled")
        h0, f0, ti0 = plan.run(n_t, fi, F, hi, 50, 1)
        plan.set_tuning(_lib.TUNE_PERSISTENT, 1)
        h1, f1, ti1 = plan.run(n_t, fi, F, hi, 50, 1)
        assert plan.stats()["launches"] <= 2  # the march and the last snapshot
        assert ti1 == ti0 > 20
        np.testing.assert_array_equal(h1, h0)
        np.testing.assert_array_equal(f1, f0)
        # fewer workgroups than row blocks: each sweeps blocks lb, lb + grid, ...
        assert K.shape[0] > 256  # two or more row blocks of 256 rows
        for cap in (1,):
            plan.set_tuning(_lib.TUNE_PERSISTENT_GRID, cap)
            hc, fc, _ = plan.run(n_t, fi, F, hi, 50, 1)
            np.testing.assert_array_equal(hc, h0)
            np.testing.assert_array_equal(fc, f0)
        plan.set_tuning(_lib.TUNE_PERSISTENT_GRID, 0)
        # a second persistent run on the same plan (buffers and lists rebuilt)
        h2, _, _ = plan.run(n_t, fi[:1], F[:1], hi, 50, 1)
        plan.set_tuning(_lib.TUNE_PERSISTENT, 0)
        h3, _, _ = plan.run(n_t, fi[:1], F[:1], hi, 50, 1)
        np.testing.assert_array_equal(h2, h3)
    finally:
        plan.close()
    r, rf, _, rti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi + 1, F, hi + 1, 50, "KE", mode)
    assert rti == ti1
    assert np.linalg.norm(r) > 0
    assert rel_l2(h1, r) < TOL
    assert rel_l2(f1, rf) < TOL


@pytest.mark.parametrize("mode", MODES)
def test_cube_3d_tets_parity(mode):
    """C3D4 cube with side-face absorbing layers, 3 x 3-node patch source,
    16 top-face monitors (the 3D workload at 1 mm elements)."""
    print("case: cube_3d tets", mode, flush=True)
    from bristolfe_v2_amd import configs
    # at half the CFL step, where both solver modes are stable (at the full
    # step dt c/m reaches 3.2 in the layers and the 'vel at last half time
    # step' recurrence diverges)
    dt = 0.5 * configs.cube_3d(el_size=1e-3, n_t=2)["dt"]
    c = configs.cube_3d(el_size=1e-3, n_t=400, dt=dt)
    K, C, M = c["K"], c["C"], c["M"]
    hi = np.concatenate([c["hist_gi"], np.arange(1, K.shape[0] + 1, 331)])
    h, _, _, _ = _gpu(K, C, M, c["time"], c["frc_gi"], c["frcs"], hi, mode=mode)
    r, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, c["time"], c["frc_gi"], c["frcs"], hi,
                                                solver_mode=mode)
    assert np.linalg.norm(r[:16]) > 0
    assert np.abs(r).max() < 1e-6   # a stable run (displacements of a few nm)
    assert rel_l2(h, r) < TOL


@pytest.mark.parametrize("operator", ["csr", "ebe"])
def test_graph_captured_march_matches_eager(operator):
    """The time loop captured in one hipGraph gives the eager march's
    histories and field frames bit for bit; repeated graph runs, with eager
    runs and synchronisations between them, stay identical (coupled model:
    interface solve in the loop; "ebe": element-by-element groups)."""
    print("case: graph capture", operator, flush=True)
    from bristolfe_v2_amd import _lib, ebe
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3)
    opts = {"dof_to_use": [1, 2, 4], "damping_power_law": 2}
    K, C, M, gl, t, F, fi, hi = _setup(mod, matls, opts, 400, 5e6, src=[1.5e-3, 2.0e-3], dof=4)
    op = ebe.element_operator(mod, matls, opts, gl, K.shape[0]) if operator == "ebe" else None
    plan = PKG.make_plan(K, C, M, t[1] - t[0], element_op=op, operator=operator)
    try:
        assert plan.info()["operator_kind"] == (1 if operator == "ebe" else 0)
        ref, rfld, ti = plan.run(t.size, fi - 1, F, hi - 1, 50, 1)
        for _ in range(3):
            plan.set_tuning(_lib.TUNE_GRAPH, 1)
            h, fld, ti2 = plan.run(t.size, fi - 1, F, hi - 1, 50, 1)
            assert ti2 == ti
            np.testing.assert_array_equal(h, ref)
            np.testing.assert_array_equal(fld, rfld)
            plan.set_tuning(_lib.TUNE_GRAPH, 0)
            h2, _, _ = plan.run(t.size, fi - 1, F, hi - 1, 50, 1)
            np.testing.assert_array_equal(h2, ref)
    finally:
        plan.close()


@pytest.mark.parametrize("prec", ["double", "single"])
def test_device_ke_field_matches_host_reduction(prec):
    """bfe_run_field_ke: the 'KE' field reduced per element on the device
    against the oracle's reduction of the oracle's velocity frames (coupled
    model: CPE3, AC2D3 and 2-node ASI2D2 rows)."""
    print("case: device KE field", prec, flush=True)
    from oracle import post as op
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3)
    opts = {"dof_to_use": [1, 2, 4], "damping_power_law": 2}
    K, C, M, gl, t, F, fi, hi = _setup(mod, matls, opts, 400, 5e6, src=[1.5e-3, 2.0e-3], dof=4)
    plan = PKG.make_plan(K, C, M, t[1] - t[0], solver_precision=prec)
    try:
        plan.set_field_elements(mod["els"], gl)
        h, el, ti = plan.run_field_ke(t.size, fi - 1, F, hi - 1, 40)
    finally:
        plan.close()
    r, rf, _, rti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, 40, "KE")
    ke = op.field_output_ke(rf, mod["els"], M.diagonal(), gl)
    assert ti == rti and el.shape == ke.shape and np.linalg.norm(ke) > 0
    tol = TOL if prec == "double" else 1e-3
    assert rel_l2(h, r) < tol
    assert rel_l2(el, ke) < tol


FIELD_TYPES = ["mean(u1)", "mean(v2)", "mean(u4)", "div(u)", "div(v)", "curl(u)", "curl(v)"]


@pytest.mark.parametrize("prec", ["double", "single"])
@pytest.mark.parametrize("ftype", FIELD_TYPES)
def test_device_element_fields_match_host_reduction(ftype, prec):
    """bfe_run_field: the 'mean', 'div' and 'curl' fields reduced per element
    on the device. Coupled model with a void: CPE3 and AC2D3 elements, and
    2-node ASI2D2 rows (an absent third node: 0 for div and curl). In FP64
    the device values equal the host restatement (field_output.py, pinned by
    closed forms) applied to the same run's frames bit for bit, and the
    oracle's frames reduced on the host to 1e-10."""
    print("case: device element field", ftype, prec, flush=True)
    from bristolfe_v2_amd import driver
    void = PKG.mesh.fn_create_smooth_random_blob(0.5, 3, 48, 7) * 0.3e-3 + np.array([1.5e-3, 0.7e-3])
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.1e-3, void=void)
    opts = {"dof_to_use": [1, 2, 4], "damping_power_law": 2}
    K, C, M, gl, t, F, fi, hi = _setup(mod, matls, opts, 400, 5e6, src=[1.5e-3, 2.0e-3], dof=4)
    kind = driver.device_field_kind(ftype, 2)
    mats = {"M": M, "gl_lookup": gl}
    plan = PKG.make_plan(K, C, M, t[1] - t[0], solver_precision=prec)
    try:
        plan.set_field_elements(mod["els"], gl)
        plan.set_field_geometry(mod["nds"])
        h, el, ti = plan.run_field(t.size, fi - 1, F, hi - 1, 40, *kind)
        h2, fld, ti2 = plan.run(t.size, fi - 1, F, hi - 1, 40, 1 if kind[1] else 0)
    finally:
        plan.close()
    host = PKG.fn_get_field_output(fld, mod, mats, ftype)
    assert ti == ti2 and el.shape == host.shape and np.linalg.norm(host) > 0
    np.testing.assert_array_equal(h, h2)
    if ftype.startswith(("div", "curl")):
        invalid = np.any(np.asarray(mod["els"]) == 0, axis=1)
        assert invalid.any() and not el[invalid].any()
    if prec == "double":
        np.testing.assert_array_equal(el, host)
    else:
        assert rel_l2(el, host) < 1e-5
    r, rf, _, rti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, 40, ftype)
    ref = PKG.fn_get_field_output(rf, mod, mats, ftype)
    assert ti == rti
    assert rel_l2(el, ref) < (TOL if prec == "double" else 1e-3)


def test_device_element_field_errors():
    """Argument checks of bfe_set_field_geometry / bfe_run_field."""
    print("case: device element field errors", flush=True)
    from bristolfe_v2_amd import _lib, configs
    mod, matls = coupled_block(3e-3, 1.5e-3, 1.0e-3, 0.25e-3)
    opts = {"dof_to_use": [1, 2, 4]}
    K, C, M, gl, t, F, fi, hi = _setup(mod, matls, opts, 50, 5e6, src=[1.5e-3, 2.0e-3], dof=4)
    with PKG.make_plan(K, C, M, t[1] - t[0]) as plan:
        with pytest.raises(_lib.BfeError):  # no element tables yet
            plan.set_field_geometry(mod["nds"])
        plan.set_field_elements(mod["els"], gl)
        with pytest.raises(_lib.BfeError, match="geometry"):
            plan.run_field(t.size, fi - 1, F, hi - 1, 10, _lib.FIELD_DIV, 0)
        with pytest.raises(_lib.BfeError, match="component"):
            plan.run_field(t.size, fi - 1, F, hi - 1, 10, _lib.FIELD_MEAN, 0, 4)
        with pytest.raises(_lib.BfeError, match="unknown field type"):
            plan.run_field(t.size, fi - 1, F, hi - 1, 10, 7, 0)
        # an all-zero forcing: v6's loop is empty (v6:163-180), every frame is 0
        h, el, ti = plan.run_field(t.size, fi - 1, np.zeros_like(F), hi - 1, 10, _lib.FIELD_MEAN,
                                   0, 1)
        assert ti == t.size + 1 and not h.any() and not el.any()
        assert el.shape == (mod["els"].shape[0], (t.size - 1) // 10 + 1)
    c = configs.cube_3d(el_size=1e-3, n_t=20)
    with PKG.make_plan(c["K"], c["C"], c["M"], c["dt"]) as plan:
        plan.set_field_elements(c["mod"]["els"], c["gl_lookup"])
        plan.set_field_geometry(c["mod"]["nds"])
        with pytest.raises(_lib.BfeError, match="curl for 3D"):
            plan.run_field(20, c["frc_gi"] - 1, c["frcs"], c["hist_gi"] - 1, 10, _lib.FIELD_CURL, 0)
        h, el, _ = plan.run_field(20, c["frc_gi"] - 1, c["frcs"], c["hist_gi"] - 1, 10,
                                  _lib.FIELD_DIV, 0)
        assert el.shape == (c["mod"]["els"].shape[0], 2) and np.isfinite(el).all()


def test_driver_device_fields_match_host_path():
    """fn_FE_entry_point with each field_output_type: the device reduction
    (default) against device_field_output = False (frames downloaded, host
    reduction), bit for bit."""
    print("case: driver device fields", flush=True)
    mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4]})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    t = np.arange(200) * dt
    nd = PKG.mesh.fn_find_node_at_point(mod["nds"], [1e-3, 1.5e-3])
    steps = {"load": {"time": t, "frcs": PKG.mesh.hann_toneburst(t, 5e6, 3), "frc_nds": [nd],
                      "frc_dfs": [4]},
             "mon": {"nds": [nd], "dfs": [4]}}
    for ftype in ["KE", "mean(u2)", "div(v)", "curl(u)"]:
        opts = {"dof_to_use": [1, 2, 4], "field_output_every_n_frames": 25,
                "field_output_type": ftype}
        dev = PKG.fn_FE_entry_point(mod, matls, steps, opts)[0]
        host = PKG.fn_FE_entry_point(mod, matls, steps, dict(opts, device_field_output=False))[0]
        assert np.linalg.norm(dev["fld"]) > 0
        np.testing.assert_array_equal(dev["fld_time"], host["fld_time"])
        if ftype == "KE":  # sqrt of a sum: the device and host orders differ in rounding
            assert rel_l2(dev["fld"], host["fld"]) < 1e-14
        else:
            np.testing.assert_array_equal(dev["fld"], host["fld"])


@pytest.mark.parametrize("mode", MODES)
def test_batched_row_groups_tiny_model(mode):
    """A 9-DOF model (one steel and one water triangle with their interface;
    fewer rows than one 64-row slice, odd row count) marched with 1 and 3
    source columns: every row group and rows-per-wavefront shape matches the
    single-source runs. (This mesh's step map has spectral radius 1.07 at the
    fixture's dt; the forced DOFs are ones whose histories are well
    conditioned: a 1e-15 change of the forcing moves them by < 1e-14.
    Forcing DOF 0 gives a 4e-10 response that rounding moves by 1e-8.)"""
    print("case: batched tiny model", mode, flush=True)
    from pathlib import Path

    import scipy.sparse as sp

    from bristolfe_v2_amd import _lib
    z = np.load(Path(__file__).resolve().parent / "golden" / "stability_mesh2.npz")
    n = z["m"].size
    K = sp.csr_matrix((z["k_val"], z["k_col"], z["k_ptr"]), shape=(n, n))
    C = sp.csr_matrix((z["c_val"], z["c_col"], z["c_ptr"]), shape=(n, n))
    M = sp.diags(z["m"])
    n_t = int(z["n_t"])
    F = np.sin(np.arange(n_t) * 0.3) * np.exp(-0.05 * np.arange(n_t))
    hi = np.arange(n)
    plan = PKG.make_plan(K, C, M, float(z["dt"]), mode)
    try:
        dofs = (n - 1, 4, 6)
        singles = [plan.run(n_t, [d], F, hi)[0] for d in dofs]
        errs = {}
        for S in (1, 3):
            for g in (1, 2, 4, 8):
                for rpw in sorted({g, 16}):
                    plan.set_tuning(_lib.TUNE_MULTI_GROUP, g)
                    plan.set_tuning(_lib.TUNE_MULTI_ROWS_PER_WAVE, rpw)
                    hb, _ = plan.run_batched(n_t, S, np.array(dofs[:S]), np.arange(S), F, hi)
                    assert plan.info()["multi_group"] == g
                    for s in range(S):
                        errs[(S, g, rpw, s)] = rel_l2(hb[:, s, :], singles[s])
        assert max(errs.values()) < 1e-13, errs
    finally:
        plan.close()

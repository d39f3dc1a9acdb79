"""Seeded random cases on the GPU against the oracle's literal v6: mesh kind
(right-triangle plate, isometric plate, fluid-solid block with interface
rows), size, absorbing layer, void, solver mode, forcing layout (one row
broadcast to several entries, or one row per entry with leading zeros and
duplicate entries, the last duplicate winning), monitors (with duplicates)
and field output; small C3D4 cubes with random side layers; jittered
(unstructured) plates, which the CSR operator marches; and the device
assembly of jittered meshes against the numpy assembler. Each case
runs the CSR operator and the element operator (FP64, 1e-10 relative L2 on
histories and field frames; FP32 within 1e-3), and the batched march (one
column per forcing entry) on the CSR operator, the multi-source element
product and its node-block form against single-source runs (1e-13) and the
oracle (1e-10)."""
import numpy as np
import pytest

from conftest import PKG, assemble, coupled_block, rect, rel_l2, steel
from oracle import v6 as ov6

pytestmark = pytest.mark.gpu

MODES = ["vel at curent time step", "vel at last half time step"]
TOL = 1e-10


def _case(seed):
    rng = np.random.default_rng(seed)
    m = PKG.mesh
    kind = ["plate", "iso", "coupled"][seed % 3]
    el = float(rng.choice([0.05e-3, 0.1e-3, 0.125e-3, 0.2e-3]))
    if seed >= 9:
        el = 0.03e-3  # tens of thousands of DOF: whole wavefronts of regular nodes
    w = float(rng.uniform(1.5e-3, 5e-3))
    if kind == "coupled":
        hs, hf = float(rng.uniform(0.8e-3, 1.6e-3)), float(rng.uniform(0.5e-3, 1.0e-3))
        void = None
        if rng.random() < 0.5:
            void = m.fn_create_smooth_random_blob(0.5, 3, 48, int(seed)) * 0.25e-3 + \
                np.array([w / 2, hs / 2])
        mod, matls = coupled_block(w, hs, hf, el, absorbing=bool(rng.random() < 0.7), void=void)
        opts = {"dof_to_use": [1, 2, 4], "damping_power_law": int(rng.choice([2, 3]))}
    else:
        h = float(rng.uniform(1e-3, 2.5e-3))
        poly = rect(w, h)
        if kind == "iso":
            mod = m.fn_isometric_structured_mesh(poly, el)
        else:
            mod = m.fn_rectangular_structured_mesh(poly, el)
        mod["el_typ_i"] = np.array(["CPE3"] * len(mod["el_mat_i"]), dtype=object)
        if rng.random() < 0.7:
            t_l = float(rng.uniform(0.2, 0.35)) * min(w, h)
            inner = np.array([[t_l, t_l], [w - t_l, t_l], [w - t_l, 10 * h], [t_l, 10 * h]])
            mod = m.fn_add_absorbing_layer(mod, inner, t_l)
        matls = [steel()]
        opts = {"dof_to_use": [1, 2], "damping_power_law": int(rng.choice([2, 3]))}
    K, C, M, gl = assemble(mod, matls, opts)
    n = K.shape[0]
    dt = m.stable_time_step(K, M.diagonal())
    if opts["damping_power_law"] and C.nnz:
        dt *= 0.5  # both solver modes stable in the layers (dt c / m stays below 2)
    n_t = int(rng.integers(150, 300))
    t = np.arange(n_t) * dt
    p = int(rng.integers(1, 5))
    fi = rng.choice(n, size=p, replace=False)
    if p > 1 and rng.random() < 0.5:
        fi[-1] = fi[0]  # a duplicate entry: the later one wins (v6:182)
    if rng.random() < 0.4:
        F = m.hann_toneburst(t, 5e6, 3)[None, :]  # one row, broadcast to every entry
    else:
        F = np.zeros((p, n_t))
        for r in range(p):
            d = int(rng.integers(0, n_t // 3))
            F[r, d:] = m.hann_toneburst(t[:n_t - d], float(rng.uniform(3e6, 6e6)), 3)
    q = int(rng.integers(3, 20))
    hi = rng.choice(n, size=q, replace=True)  # duplicates allowed
    every = int(rng.choice([0, 17, 40]))
    mode = MODES[int(rng.integers(0, 2))]
    return kind, mod, matls, opts, K, C, M, gl, t, F, fi, hi, every, mode


@pytest.mark.parametrize("seed", range(12))
def test_random_case_against_oracle(seed):
    from bristolfe_v2_amd import ebe, solver
    kind, mod, matls, opts, K, C, M, gl, t, F, fi, hi, every, mode = _case(seed)
    print(f"case: random seed {seed} {kind} n={K.shape[0]} n_t={t.size} p={fi.size} "
          f"rows={F.shape[0]} q={hi.size} every={every} {mode}", flush=True)
    dt = t[1] - t[0]
    r, rf, rft, rti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi + 1, F, hi + 1,
                                                      every if every else np.inf, "KE", mode)
    assert np.isfinite(r).all() and np.linalg.norm(r) > 0
    op = ebe.element_operator(mod, matls, opts, gl, K.shape[0])
    runs = {}
    with solver.make_plan(K, C, M, dt, mode, operator="csr") as plan:
        runs["csr"] = plan.run(t.size, fi, F, hi, every, 1)
        # the whole loop as one hipGraph and as one persistent launch: bit-identical
        from bristolfe_v2_amd import _lib
        for key in (_lib.TUNE_GRAPH, _lib.TUNE_PERSISTENT):
            plan.set_tuning(key, 1)
            hg, fg, tg = plan.run(t.size, fi, F, hi, every, 1)
            plan.set_tuning(key, 0)
            np.testing.assert_array_equal(hg, runs["csr"][0])
            np.testing.assert_array_equal(fg, runs["csr"][1])
            assert tg == runs["csr"][2]
        # batched: one column per forcing entry (the entry's own signal row)
        rows = F if F.shape[0] > 1 else np.repeat(F, fi.size, axis=0)
        hb, tib = plan.run_batched(t.size, fi.size, fi, np.arange(fi.size), rows, hi)
        for s in range(fi.size):
            h1, _, _ = plan.run(t.size, fi[s:s + 1], rows[s:s + 1], hi)
            assert rel_l2(hb[:, s, :], h1) < 1e-13
            r1, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi[s:s + 1] + 1, rows[s:s + 1],
                                                         hi + 1, np.inf, "KE", mode)
            assert rel_l2(hb[:, s, :], r1) < TOL
    if op is not None:
        with solver.make_plan(K, C, M, dt, mode, element_op=op, operator="ebe") as plan:
            runs["ebe"] = plan.run(t.size, fi, F, hi, every, 1)
            _batched_operators(plan, t, fi, rows, hi, hb)
        with solver.make_plan(K, C, M, dt, mode, "single", element_op=op, operator="ebe") as plan:
            h32, _, _ = plan.run(t.size, fi, F, hi)
        print(f"  fp32 element operator: rel L2 {rel_l2(h32, r):.2e}", flush=True)
        assert rel_l2(h32, r) < 1e-3
    for name, (h, fld, ti) in runs.items():
        assert ti == rti, name
        err = rel_l2(h, r)
        print(f"  {name}: history rel L2 {err:.2e}", flush=True)
        assert err < TOL, name
        if every:
            assert rel_l2(fld, rf) < TOL, name


def _batched_operators(plan, t, fi, rows, hi, h_csr):
    """The multi-source element product and the node-block form on the same
    columns: both within 1e-13 of the CSR batched march (the node-block form
    where the mesh pattern has it, CSR otherwise)."""
    from bristolfe_v2_amd import _lib
    S = fi.size
    for sel in (_lib.OPERATOR_EBE, _lib.OPERATOR_NODE):
        plan.set_tuning(_lib.TUNE_MULTI_OPERATOR, sel)
        h, _ = plan.run_batched(t.size, S, fi, np.arange(S), rows, hi)
        used = plan.info()["multi_operator"]
        assert used in (sel, _lib.OPERATOR_CSR)
        assert rel_l2(h, h_csr) < 1e-13, (sel, used)
    plan.set_tuning(_lib.TUNE_MULTI_OPERATOR, _lib.OPERATOR_CSR)


@pytest.mark.parametrize("seed", range(3))
def test_random_tet_cube_against_oracle(seed):
    """Small C3D4 cubes: random size, element size, side-layer thickness,
    mode and forcing on the top-face patch, at half the stable step (both
    modes stable in the layers)."""
    from bristolfe_v2_amd import configs, ebe, solver
    rng = np.random.default_rng(100 + seed)
    # (size, element size) pairs whose cell count is exact in floating point
    size, el = [(3e-3, 0.25e-3), (2.5e-3, 0.25e-3), (3.5e-3, 0.25e-3), (4e-3, 0.5e-3)][
        int(rng.integers(0, 4))]
    layer = float(rng.uniform(0.2, 0.3)) * size
    mode = MODES[int(rng.integers(0, 2))]
    n_t = int(rng.integers(100, 180))
    c0 = configs.cube_3d(size=size, el_size=el, layer=layer, n_t=2, n_mon=4, mon_pitch=el)
    c = configs.cube_3d(size=size, el_size=el, layer=layer, n_t=n_t, n_mon=4, mon_pitch=el,
                        dt=0.5 * c0["dt"])
    K, C, M, t = c["K"], c["C"], c["M"], c["time"]
    fi = c["frc_gi"] - 1
    hi = np.unique(np.concatenate([c["hist_gi"] - 1, rng.choice(K.shape[0], 12)]))
    print(f"case: random tet cube seed {seed} n={K.shape[0]} size={size:.2e} el={el:.2e} "
          f"{mode}", flush=True)
    r, _, _, rti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi + 1, c["frcs"], hi + 1,
                                                   np.inf, "KE", mode)
    assert np.isfinite(r).all() and np.linalg.norm(r) > 0
    op = ebe.element_operator(c["mod"], c["matls"], c["fe_options"], c["gl_lookup"], K.shape[0])
    assert op is not None
    for operator in ("csr", "ebe"):
        with solver.make_plan(K, C, M, c["dt"], mode, element_op=op, operator=operator) as plan:
            h, _, ti = plan.run(n_t, fi, c["frcs"], hi)
        err = rel_l2(h, r)
        print(f"  {operator}: rel L2 {err:.2e}", flush=True)
        assert ti == rti and err < TOL


@pytest.mark.parametrize("seed", range(2))
def test_random_unstructured_mesh(seed):
    """A plate whose nodes are jittered (no two elements alike): the element
    operator does not apply, make_plan("auto") marches the CSR operator, and
    the histories match the oracle."""
    from bristolfe_v2_amd import ebe, solver
    rng = np.random.default_rng(300 + seed)
    m = PKG.mesh
    w, h, el = 3e-3, 1.5e-3, 0.1e-3
    mod = m.fn_rectangular_structured_mesh(rect(w, h), el)
    mod["nds"] = mod["nds"] + rng.uniform(-0.25, 0.25, mod["nds"].shape) * el
    mod["el_typ_i"] = np.array(["CPE3"] * len(mod["el_mat_i"]), dtype=object)
    inner = np.array([[0.6e-3, 0.4e-3], [2.4e-3, 0.4e-3], [2.4e-3, 5e-3], [0.6e-3, 5e-3]])
    mod = m.fn_add_absorbing_layer(mod, inner, 0.4e-3)
    opts = {"dof_to_use": [1, 2], "damping_power_law": 2}
    K, C, M, gl = assemble(mod, [steel()], opts)
    assert ebe.element_operator(mod, [steel()], opts, gl, K.shape[0]) is None
    dt = 0.5 * m.stable_time_step(K, M.diagonal())
    t = np.arange(250) * dt
    mode = MODES[seed % 2]
    fi = rng.choice(K.shape[0], 2, replace=False)
    F = m.hann_toneburst(t, 5e6, 3)[None, :]
    hi = rng.choice(K.shape[0], 15, replace=False)
    print(f"case: random unstructured seed {seed} n={K.shape[0]} {mode}", flush=True)
    with solver.make_plan(K, C, M, dt, mode, operator="auto") as plan:
        assert plan.info()["operator_kind"] == 0
        h, fld, ti = plan.run(t.size, fi, F, hi, 50, 1)
    r, rf, _, rti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi + 1, F, hi + 1, 50, "KE", mode)
    assert ti == rti and np.linalg.norm(r) > 0
    assert rel_l2(h, r) < TOL and rel_l2(fld, rf) < TOL


@pytest.mark.parametrize("seed", range(8))
def test_random_device_assembly(seed):
    """The device assembly (bfe_assemble_device) against the numpy assembler
    on jittered (unstructured) versions of the random meshes above and of
    small C3D4 cubes with random absorption: identical DOF numbering, values
    to 1e-13."""
    rng = np.random.default_rng(400 + seed)
    if seed < 6:
        kind, mod, matls, opts, *_ = _case(seed)
        mod = dict(mod)
        el = float(np.min(np.abs(np.diff(np.unique(np.round(mod["nds"][:, 0], 12))))))
    else:
        kind, el = "tets", 0.25e-3
        mod = PKG.mesh.cubic_structured_tet_mesh(float(rng.choice([0.75e-3, 1e-3])), el)
        mod["el_typ_i"] = np.array(["C3D4"] * len(mod["el_mat_i"]), dtype=object)
        mod["el_abs_i"] = rng.uniform(0, 1, len(mod["el_mat_i"])) * (rng.random(len(mod["el_mat_i"])) < 0.5)
        matls = [steel()]
        opts = {"dof_to_use": [] if seed == 6 else [1, 2, 3], "damping_power_law": 2}
    # each coordinate moved by up to a fifth of the smallest spacing: no
    # element turns over (checked on the host below)
    mod["nds"] = mod["nds"] + rng.uniform(-0.2, 0.2, mod["nds"].shape) * el
    if kind == "tets":
        e = np.asarray(mod["els"], np.int64)[:, :4] - 1
        x = mod["nds"]
        vol = np.einsum("ij,ij->i", np.cross(x[e[:, 1]] - x[e[:, 0]], x[e[:, 2]] - x[e[:, 0]]),
                        x[e[:, 3]] - x[e[:, 0]])
        assert np.all(vol != 0) and len(np.unique(np.sign(vol))) == 1
    print(f"case: random device assembly seed {seed} {kind} n_el={len(mod['el_mat_i'])}", flush=True)
    K1, C1, M1, gl1 = assemble(mod, matls, dict(opts, assembly_backend="numpy"))
    K2, C2, M2, gl2 = assemble(mod, matls, dict(opts, assembly_backend="device"))
    np.testing.assert_array_equal(gl1, gl2)
    for A, B in ((K1, K2), (C1, C2), (M1, M2)):
        assert A.shape == B.shape
        scale = max(abs(A).max(), abs(B).max(), 1e-300)
        assert abs(A - B).max() <= 1e-13 * scale

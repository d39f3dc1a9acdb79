"""Differentiation: per-element stiffness-scale and density gradients of an
A-scan misfit. CPU: the parametrised matrices against the oracle's assembly
with per-element materials. GPU: the adjoint gradients against central finite
differences of the oracle's literal v6 march."""
import numpy as np
import pytest

from conftest import PKG, coupled_block, rect, rel_l2, solid_plate, steel
from oracle import assembly as oa
from oracle import v6 as ov6

MODES = ["vel at curent time step", "vel at last half time step"]


def _adj():
    from bristolfe_v2_amd import adjoint
    return adjoint


def _small_solid(absorbing=True):
    m = PKG.mesh
    mod, matls = solid_plate(2e-3, 1e-3, 0.125e-3)
    if absorbing:
        inner = np.array([[0.4e-3, 0.3e-3], [1.6e-3, 0.3e-3], [1.6e-3, 5e-3], [0.4e-3, 5e-3]])
        mod = m.fn_add_absorbing_layer(mod, inner, 0.3e-3)
    return mod, matls


def test_parametrized_matrices_match_per_element_materials():
    adj = _adj()
    mod, matls = _small_solid()
    opts = {"dof_to_use": [1, 2], "damping_power_law": 2, "max_damping": 2e7}
    pm = adj.ParametrizedModel(mod, matls, opts)
    rng = np.random.default_rng(3)
    s = 1 + 0.3 * rng.standard_normal(pm.n_el)
    rho = pm.rho0 * (1 + 0.2 * rng.standard_normal(pm.n_el))
    K, C, M = pm.matrices(s, rho)
    # oracle: one material per element with D*s_e and rho_e
    n_el = len(mod["el_mat_i"])
    mats = []
    s_full = np.ones(n_el)
    r_full = np.full(n_el, 8000.0)
    s_full[pm.el_index] = s
    r_full[pm.el_index] = rho
    for e in range(n_el):
        mats.append({"rho": r_full[e], "D": steel()["D"] * s_full[e], "el_typ": "CPE3"})
    K2, C2, M2, gl2 = oa.build_global_matrices(mod["nds"], mod["els"], np.arange(1, n_el + 1),
                                               mod["el_abs_i"], ["CPE3"] * n_el, mats, opts)
    np.testing.assert_array_equal(pm.gl, gl2)
    for A, B in ((K, K2), (C, C2), (M, M2)):
        assert abs(A - B).max() <= 1e-12 * abs(B).max()


def _misfit_oracle(pm, shots, data, s, rho, mode):
    K, C, M = pm.matrices(s, rho)
    t = np.arange(shots.n_t) * shots.dt
    J = 0.0
    for i, tx in enumerate(shots.tx):
        h, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, tx + 1, shots.signal, shots.rx + 1,
                                                    solver_mode=mode)
        J += 0.5 * np.sum((h - data[i]) ** 2)
    return J


def _case(kind, mode):
    adj = _adj()
    if kind == "solid":
        mod, matls = _small_solid()
        opts = {"dof_to_use": [1, 2], "damping_power_law": 2, "max_damping": 2e7, "solver_mode": mode}
    else:
        mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
        opts = {"dof_to_use": [1, 2, 4], "damping_power_law": 2, "max_damping": 2e7,
                "solver_mode": mode}
    pm = adj.ParametrizedModel(mod, matls, opts)
    dt = PKG.mesh.stable_time_step(pm.K0, pm.M0.diagonal())
    n_t = 260
    sig = PKG.mesh.hann_toneburst(np.arange(n_t) * dt, 6e6, 2)
    gl = pm.gl
    nds = mod["nds"]
    top = np.nonzero(np.isclose(nds[:, 1], nds[:, 1][gl[:, 1] > 0].max()) & (gl[:, 1] > 0))[0]
    top = top[np.argsort(nds[top, 0])]
    tx_nodes = top[[len(top) // 4, len(top) // 2, 3 * len(top) // 4]]
    tx = [np.array([gl[k, 1] - 1]) for k in tx_nodes]
    rx = np.array([gl[k, 1] - 1 for k in top[::2]])
    shots = adj.Shots(tx, sig, rx, n_t, dt, ckpt_every=37)
    # synthetic data: a perturbed model
    rng = np.random.default_rng(11)
    s_true = 1 + 0.05 * rng.standard_normal(pm.n_el)
    r_true = pm.rho0 * (1 + 0.03 * rng.standard_normal(pm.n_el))
    K, C, M = pm.matrices(s_true, r_true)
    t = np.arange(n_t) * dt
    data = np.stack([ov6.explicit_dynamic_solver_v6(K, C, M, t, x + 1, sig, rx + 1, solver_mode=mode)[0]
                     for x in tx])
    return pm, shots, data


@pytest.mark.gpu
@pytest.mark.parametrize("kind,mode", [("solid", MODES[0]), ("solid", MODES[1]), ("coupled", MODES[0])])
def test_adjoint_gradient_matches_finite_differences(kind, mode):
    print("case: adjoint", kind, mode, flush=True)
    adj = _adj()
    pm, shots, data = _case(kind, mode)
    s0 = np.ones(pm.n_el)
    r0 = pm.rho0.copy()
    A, plan, _ = adj.forward(pm, shots, s0, r0, keep=True)
    # the GPU forward agrees with the oracle
    K, C, M = pm.matrices(s0, r0)
    t = np.arange(shots.n_t) * shots.dt
    ref = ov6.explicit_dynamic_solver_v6(K, C, M, t, shots.tx[0] + 1, shots.signal, shots.rx + 1,
                                         solver_mode=mode)[0]
    assert rel_l2(A[0], ref) < 1e-10
    gk, gm, _ = adj.backward(plan, A - data)
    plan.close()
    # central differences of the oracle misfit on the most sensitive elements
    rng = np.random.default_rng(5)
    cand = np.argsort(-np.abs(gk))[:40]
    pick = rng.choice(cand, 4, replace=False)
    for e in pick:
        for which, g, base in (("s", gk, s0), ("rho", gm, r0)):
            h = 1e-4 * abs(base[e])
            p, m = base.copy(), base.copy()
            p[e] += h
            m[e] -= h
            if which == "s":
                Jp = _misfit_oracle(pm, shots, data, p, r0, mode)
                Jm = _misfit_oracle(pm, shots, data, m, r0, mode)
            else:
                Jp = _misfit_oracle(pm, shots, data, s0, p, mode)
                Jm = _misfit_oracle(pm, shots, data, s0, m, mode)
            fd = (Jp - Jm) / (2 * h)
            err = abs(g[e] - fd) / max(abs(fd), 1e-300)
            print(f"  {which}[{e}] adjoint {g[e]:.6e} fd {fd:.6e} rel {err:.2e}")
            assert err < 1e-5


@pytest.mark.gpu
def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    adj = _adj()
    pm, shots, data = _case("solid", MODES[0])
    s = torch.ones(pm.n_el, dtype=torch.float64, requires_grad=True)
    rho = torch.tensor(pm.rho0, dtype=torch.float64, requires_grad=True)
    J = adj.ascan_misfit(pm, shots, data, s, rho)
    J.backward()
    A, plan, _ = adj.forward(pm, shots, np.ones(pm.n_el), pm.rho0, keep=True)
    gk, gm, _ = adj.backward(plan, A - data)
    plan.close()
    # the wrapper marches all transmitters in one batched run: the same
    # misfit and gradients as the per-transmitter runs, to rounding
    assert np.isclose(J.item(), 0.5 * np.sum((A - data) ** 2), rtol=1e-11)
    assert rel_l2(s.grad.numpy(), gk) < 1e-11
    assert rel_l2(rho.grad.numpy(), gm) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("kind,mode", [("solid", MODES[0]), ("solid", MODES[1]), ("coupled", MODES[0])])
def test_batched_adjoint_matches_per_transmitter_adjoint(kind, mode):
    """bfe_run_checkpointed_batched + bfe_adjoint (all transmitters in one
    2S-column reverse sweep, gradients summed over the sources by
    k_grad_multi) against the sum of the per-transmitter adjoints, and its
    A-scans against the oracle; a 37-step checkpoint interval leaves a
    partial first segment. Both solver modes; the coupled model carries
    interface rows."""
    print("case: batched adjoint", kind, mode, flush=True)
    adj = _adj()
    pm, shots, data = _case(kind, mode)
    s0 = np.ones(pm.n_el)
    r0 = pm.rho0.copy()
    A1, plan1, _ = adj.forward(pm, shots, s0, r0, keep=True)
    gk1, gm1, _ = adj.backward(plan1, A1 - data)
    plan1.close()
    A2, plan2, _ = adj.forward_batched(pm, shots, s0, r0, keep=True)
    gk2, gm2, _ = adj.backward_batched(plan2, A2 - data)
    plan2.close()
    assert rel_l2(A2, A1) < 1e-12
    K, C, M = pm.matrices(s0, r0)
    t = np.arange(shots.n_t) * shots.dt
    for i in (0, len(shots.tx) - 1):
        ref = ov6.explicit_dynamic_solver_v6(K, C, M, t, shots.tx[i] + 1, shots.signal,
                                             shots.rx + 1, solver_mode=mode)[0]
        assert rel_l2(A2[i], ref) < 1e-10
    assert np.linalg.norm(gk1) > 0 and np.linalg.norm(gm1) > 0
    assert rel_l2(gk2, gk1) < 1e-11
    assert rel_l2(gm2, gm1) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("kind,mode", [("solid", MODES[0]), ("solid", MODES[1]), ("coupled", MODES[0])])
def test_batched_adjoint_node_pairs_match_element_sums(kind, mode):
    """The batched adjoint's node-pair gradient (correlations of adjoint and
    forward states on the element graph's node pairs, contracted with K_e once
    after the sweep) against the per-element sums of k_grad_multi on the same
    checkpointed run: agreement to rounding. The solid's vertices are nodes of
    two DOFs (the node-pair form applies); where a model's elements do not
    have that form (the coupled model's mixed element types) both settings
    take the per-element sums and agree exactly."""
    print("case: batched adjoint node pairs", kind, mode, flush=True)
    from bristolfe_v2_amd import _lib
    adj = _adj()
    pm, shots, data = _case(kind, mode)
    s0 = np.ones(pm.n_el)
    r0 = pm.rho0.copy()
    out = {}
    for pairs in (1, 0):
        A, plan, _ = adj.forward_batched(pm, shots, s0, r0, keep=True)
        try:
            plan.set_tuning(_lib.TUNE_ADJOINT_PAIRS, pairs)
            gp = plan.info()["grad_pairs"]
            out[pairs] = adj.backward_batched(plan, A - data)[:2]
        finally:
            plan.close()
    gk1, gm1 = out[1]
    gk0, gm0 = out[0]
    assert np.linalg.norm(gk0) > 0 and np.linalg.norm(gm0) > 0
    if kind == "solid":
        assert gp == 2
    if gp > 0:
        assert rel_l2(gk1, gk0) < 1e-12
        assert rel_l2(gm1, gm0) < 1e-12
    else:
        np.testing.assert_array_equal(gk1, gk0)
        np.testing.assert_array_equal(gm1, gm0)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_random_batched_adjoint_against_finite_differences(seed):
    """Seeded random gradient problems: solid or fluid-solid model, solver
    mode, true-model perturbation, checkpoint interval (partial first
    segments included), transmitters and receivers. The batched adjoint's
    stiffness-scale and density gradients against central differences of
    the oracle's misfit on two sensitive elements (1e-5), and its A-scans
    against the oracle (1e-10)."""
    adj = _adj()
    rng = np.random.default_rng(200 + seed)
    kind = ["solid", "coupled"][seed % 2]
    mode = MODES[(seed // 2) % 2]  # every kind in both modes
    if kind == "solid":
        mod, matls = _small_solid(absorbing=bool(rng.random() < 0.7))
        opts = {"dof_to_use": [1, 2], "damping_power_law": 2, "max_damping": 2e7, "solver_mode": mode}
    else:
        mod, matls = coupled_block(2e-3, 1.0e-3, 0.75e-3, 0.125e-3)
        opts = {"dof_to_use": [1, 2, 4], "damping_power_law": 2, "max_damping": 2e7,
                "solver_mode": mode}
    pm = adj.ParametrizedModel(mod, matls, opts)
    dt = 0.8 * PKG.mesh.stable_time_step(pm.K0, pm.M0.diagonal())
    n_t = int(rng.integers(150, 260))
    sig = PKG.mesh.hann_toneburst(np.arange(n_t) * dt, float(rng.uniform(4e6, 7e6)), 2)
    gl, nds = pm.gl, mod["nds"]
    top = np.nonzero(np.isclose(nds[:, 1], nds[:, 1][gl[:, 1] > 0].max()) & (gl[:, 1] > 0))[0]
    top = top[np.argsort(nds[top, 0])]
    n_tx = int(rng.integers(2, 4))
    tx_nodes = rng.choice(top[1:-1], n_tx, replace=False)
    tx = [np.array([gl[k, 1] - 1]) for k in tx_nodes]
    rx = np.array([gl[k, 1] - 1 for k in rng.choice(top, int(rng.integers(3, len(top))), replace=False)])
    shots = adj.Shots(tx, sig, rx, n_t, dt, ckpt_every=int(rng.integers(7, 61)))
    print(f"case: random adjoint seed {seed} {kind} {mode} n_t={n_t} tx={n_tx} rx={rx.size} "
          f"ckpt={shots.ckpt_every}", flush=True)
    s_true = 1 + 0.05 * rng.standard_normal(pm.n_el)
    r_true = pm.rho0 * (1 + 0.03 * rng.standard_normal(pm.n_el))
    K, C, M = pm.matrices(s_true, r_true)
    t = np.arange(n_t) * dt
    data = np.stack([ov6.explicit_dynamic_solver_v6(K, C, M, t, x + 1, sig, rx + 1, solver_mode=mode)[0]
                     for x in tx])
    s0, r0 = np.ones(pm.n_el), pm.rho0.copy()
    A, plan, _ = adj.forward_batched(pm, shots, s0, r0, keep=True)
    gk, gm, _ = adj.backward_batched(plan, A - data)
    plan.close()
    K0, C0, M0 = pm.matrices(s0, r0)
    ref = ov6.explicit_dynamic_solver_v6(K0, C0, M0, t, tx[-1] + 1, sig, rx + 1, solver_mode=mode)[0]
    assert rel_l2(A[-1], ref) < 1e-10
    cand = np.argsort(-np.abs(gk))[:30]
    for e in rng.choice(cand, 2, replace=False):
        for which, g, base in (("s", gk, s0), ("rho", gm, r0)):
            h = 1e-4 * abs(base[e])
            p, m = base.copy(), base.copy()
            p[e] += h
            m[e] -= h
            if which == "s":
                fd = (_misfit_oracle(pm, shots, data, p, r0, mode) -
                      _misfit_oracle(pm, shots, data, m, r0, mode)) / (2 * h)
            else:
                fd = (_misfit_oracle(pm, shots, data, s0, p, mode) -
                      _misfit_oracle(pm, shots, data, s0, m, mode)) / (2 * h)
            err = abs(g[e] - fd) / max(abs(fd), 1e-300)
            print(f"  {which}[{e}] adjoint {g[e]:.6e} fd {fd:.6e} rel {err:.2e}", flush=True)
            assert err < 1e-5

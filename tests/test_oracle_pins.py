"""Pins of the oracle (CPU): closed forms, brute force and invariants that a
plausible mistake in the restatement (dropped term, wrong sign, wrong index,
transposed operand) would fail. The reference ships no golden vectors, so
these checks are what ties the oracle to the mathematics of
fn_explicit_dynamic_solver_v6.m and the element files."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import PKG, assemble, coupled_block, rel_l2, solid_plate, steel, water
from oracle import assembly as oa
from oracle import elements as oe
from oracle import march as om
from oracle import post as op
from oracle import v6 as ov6

MODES = ["vel at curent time step", "vel at last half time step"]


# ---------------------------------------------------------------------------
# recurrence


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("c", [0.0, 3.0e4])
def test_sdof_closed_form(mode, c):
    """n = 1: u_n = u_1 r^(n-1) sin(n th)/sin(th) after a unit impulse."""
    k, m, dt = 4.0e10, 2.0, 5.0e-6
    K, C, M = sp.csr_matrix([[k]]), sp.csr_matrix([[c]]), sp.csr_matrix([[m]])
    n_t = 400
    t = np.arange(n_t) * dt
    F = np.zeros(n_t)
    F[0] = 1.0
    hist, _, _, ti = ov6.explicit_dynamic_solver_v6(K, C, M, t, [1], F, [1], solver_mode=mode)
    assert ti == 1
    if mode == MODES[0]:
        a0 = m + dt * c / 2
        u1 = dt ** 2 / a0
        r = np.sqrt((m - dt * c / 2) / a0)
        cos_th = (2 * m - dt ** 2 * k) / a0 / (2 * r)
    else:
        u1 = dt ** 2 / m
        r = np.sqrt(1 - dt * c / m)
        cos_th = (2 - dt * c / m - dt ** 2 * k / m) / (2 * r)
    th = np.arccos(cos_th)
    n = np.arange(1, n_t + 1)
    exact = u1 * r ** (n - 1) * np.sin(n * th) / np.sin(th)
    assert rel_l2(hist[0], exact) < 1e-11


def _dense_march(K, C, M, dt, n_t, fi, F, hi, mode):
    """Brute force: dense A, B, D with numpy.linalg.solve (v6:130-139)."""
    K, C, M = K.toarray(), C.toarray(), M.toarray()
    n = K.shape[0]
    m = M.sum(axis=0)
    iM = np.diag(1 / m)
    I = np.eye(n)
    if mode == MODES[1]:
        A, B, D = dt ** 2 * iM, 2 * I - dt * iM @ C - dt ** 2 * iM @ K, dt * iM @ C - I
    else:
        X = I + dt / 2 * iM @ C
        A = np.linalg.solve(X, dt ** 2 * iM)
        B = np.linalg.solve(X, 2 * I - dt ** 2 * iM @ K)
        D = np.linalg.solve(X, -I + dt / 2 * iM @ C)
    u1 = np.zeros(n)
    u2 = np.zeros(n)
    f = np.zeros(n)
    hist = np.zeros((len(hi), n_t))
    for ti in range(n_t):
        f[fi] = F[ti]
        u = A @ f + B @ u1 + D @ u2
        hist[:, ti] = u[hi]
        u2, u1 = u1, u
    return hist


def _stability_meshes():
    """The tiny meshes of tests/stability.m:41-58 (single steel element, single
    water element, pair with an interface) at el_size 1 mm."""
    h = 1e-3
    s3 = np.sqrt(3)
    m1 = {"nds": np.array([[0.5, s3 / 2], [0, 0], [1, 0]]) * h, "els": np.array([[1, 2, 3]]),
          "el_mat_i": np.array([1]), "el_abs_i": np.zeros(1), "el_typ_i": np.array(["CPE3"], dtype=object)}
    m2 = {"nds": np.array([[0, 0], [1, 0], [0.5, -s3 / 2]]) * h, "els": np.array([[1, 2, 3]]),
          "el_mat_i": np.array([2]), "el_abs_i": np.zeros(1), "el_typ_i": np.array(["AC2D3"], dtype=object)}
    m3 = {"nds": np.array([[0.5, s3 / 2], [0, 0], [1, 0], [0.5, -s3 / 2]]) * h,
          "els": np.array([[1, 2, 3], [2, 3, 4]]), "el_mat_i": np.array([1, 2]), "el_abs_i": np.zeros(2)}
    return [m1, m2, m3]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("case", [0, 1, 2])
def test_dense_brute_force_stability_meshes(pkg, mode, case):
    matls = [steel(), water()]
    mod = _stability_meshes()[case]
    if case == 2:
        mod["el_typ_i"] = np.array(["CPE3", "AC2D3"], dtype=object)
        mod = pkg.mesh.fn_add_fluid_solid_interface_els(mod, matls)
    K, C, M, gl = oa.build_global_matrices(mod["nds"], mod["els"], mod["el_mat_i"], mod["el_abs_i"],
                                           list(mod["el_typ_i"]), matls, {"dof_to_use": [1, 2, 4]})
    n = K.shape[0]
    m = M.diagonal()
    lam = np.max(np.abs(np.linalg.eigvals(np.diag(1 / m) @ K.toarray())))
    dt = 0.5 * 2 / np.sqrt(lam)
    n_t = 60
    t = np.arange(n_t) * dt
    F = np.sin(np.arange(n_t) * 0.3) * np.exp(-0.05 * np.arange(n_t))
    fi = [n]  # last DOF (pressure at the water node / a solid DOF)
    hi = np.arange(1, n + 1)
    h_or, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, solver_mode=mode)
    h_bf = _dense_march(K, C, M, dt, n_t, [n - 1], F, hi - 1, mode)
    assert rel_l2(h_or, h_bf) < 1e-12
    if case == 2:
        assert C.nnz > 0 and (C - sp.diags(C.diagonal())).nnz > 0  # interface present


def test_modes_equal_without_damping():
    mod, matls = solid_plate(1e-3, 1e-3, 0.25e-3)
    K, C, M, _ = assemble(mod, matls, {"dof_to_use": [1, 2]})
    assert C.nnz == 0
    dt = 1e-8
    A0, B0, D0 = ov6.abd_matrices(K, C, M, dt, MODES[0])
    A1, B1, D1 = ov6.abd_matrices(K, C, M, dt, MODES[1])
    for X, Y in ((A0, A1), (B0, B1), (D0, D1)):
        assert abs(X - Y).max() <= 1e-15 * max(abs(Y).max(), 1e-300)


@pytest.mark.parametrize("mode", MODES)
def test_reciprocity_coupled(mode):
    """hist_j(source i) == hist_i(source j): K, C, M are symmetric."""
    mod, matls = coupled_block(2e-3, 1e-3, 0.75e-3, 0.25e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    m = M.diagonal()
    from conftest import PKG
    dt = PKG.mesh.stable_time_step(K, m)
    n_t = 300
    t = np.arange(n_t) * dt
    F = PKG.mesh.hann_toneburst(t, 5e6, 3)
    i = int(gl[PKG.mesh.fn_find_node_at_point(mod["nds"], [1e-3, 1.5e-3]) - 1, 3])   # fluid p
    j = int(gl[PKG.mesh.fn_find_node_at_point(mod["nds"], [0.5e-3, 0.5e-3]) - 1, 1])  # solid uy
    hij, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [i], F, [j], solver_mode=mode)
    hji, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [j], F, [i], solver_mode=mode)
    assert np.linalg.norm(hij) > 0
    assert rel_l2(hij, hji) < 1e-10


def test_energy_conservation_solid():
    """1/2 v' M v (half step) + 1/2 u_{n+1}' K u_n is constant once f = 0, C = 0."""
    mod, matls = solid_plate(2e-3, 1e-3, 0.25e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2]})
    from conftest import PKG
    m = M.diagonal()
    dt = PKG.mesh.stable_time_step(K, m)
    n_t = 400
    t = np.arange(n_t) * dt
    F = PKG.mesh.hann_toneburst(t, 10e6, 2)
    assert F[150:].max() == 0
    fi = [int(gl[5, 1])]
    _, U, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, [1], 1, "mean(u1)")
    E = []
    for n in range(200, n_t - 1):
        v = (U[:, n + 1] - U[:, n]) / dt
        E.append(0.5 * v @ (m * v) + 0.5 * U[:, n + 1] @ (K @ U[:, n]))
    E = np.array(E)
    assert E[0] > 0
    assert np.max(np.abs(E - E[0])) / E[0] < 1e-10


def test_ti_start_shift_and_linearity():
    """v6:163-166: ti_start = min(find(sum(abs(F)))). MATLAB's sum of a 1 x n_t
    row is one scalar, so a single (broadcast) signal starts the march at
    ti = 1 even when its first samples are zero; with p > 1 rows the sum runs
    down the columns and the march starts at the first non-zero column."""
    mod, matls = solid_plate(1e-3, 1e-3, 0.25e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2]})
    dt = 2e-9
    n_t = 120
    t = np.arange(n_t) * dt
    F = np.zeros(n_t)
    F[:40] = np.sin(np.arange(40) * 0.4)
    G = np.concatenate([np.zeros(17), F[:-17]])
    fi, hi = [3], [3, 7, 11]
    h1, _, _, t1 = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi)
    h2, f2, ft2, t2 = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, G, hi, 5)
    h3, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, 2 * F, hi)
    assert t1 == 1 and t2 == 1    # one row: the scalar sum is non-zero
    assert np.all(h2[:, :18] == 0)
    np.testing.assert_array_equal(h2[:, 17:], h1[:, :n_t - 17])
    np.testing.assert_array_equal(h3, 2 * h1)
    # every frame is written, those before the first non-zero sample with zero
    # velocity and their own time
    frames = np.arange(1, n_t + 1, 5)
    np.testing.assert_array_equal(ft2, t[frames - 1])
    assert not f2[:, frames <= 18].any() and f2[:, frames > 19].any()
    # two rows (the second one all zero): the first column with a non-zero entry
    G2 = np.vstack([G, np.zeros(n_t)])
    h4, f4, ft4, t4 = ov6.explicit_dynamic_solver_v6(K, C, M, t, [3, 8], G2, hi, 5)
    assert t4 == 19
    np.testing.assert_array_equal(h4, h2)
    np.testing.assert_array_equal(ft4, np.where(frames >= 19, t[frames - 1], 0.0))
    np.testing.assert_array_equal(f4[:, frames >= 19], f2[:, frames >= 19])
    # NaN counts as non-zero in find(), for one row and for several
    Gn = G.copy()
    Gn[30] = np.nan
    assert ov6.ti_start_of(Gn) == 1
    assert ov6.ti_start_of(np.vstack([Gn, np.zeros(n_t)])) == 19
    Z = np.zeros((2, n_t))
    Z[1, 30] = np.nan
    assert ov6.ti_start_of(Z) == 31
    assert ov6.ti_start_of(np.zeros(n_t)) is None


def test_all_zero_forcing_gives_zero_outputs():
    K = sp.csr_matrix([[2.0, -1.0], [-1.0, 2.0]])
    M = sp.eye(2, format="csr")
    C = sp.csr_matrix((2, 2))
    t = np.arange(10) * 0.1
    h, f, ft, ti = ov6.explicit_dynamic_solver_v6(K, C, M, t, [1], np.zeros(10), [1, 2], 3)
    assert ti == 11 and not h.any() and not f.any() and not ft.any()


def test_forcing_duplicates_last_wins_and_broadcast():
    K = sp.csr_matrix([[2.0, -1.0, 0], [-1.0, 2.0, -1.0], [0, -1.0, 2.0]])
    M = sp.eye(3, format="csr")
    C = sp.csr_matrix((3, 3))
    t = np.arange(20) * 0.1
    F2 = np.vstack([np.ones(20), 3 * np.ones(20)])
    h_dup, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [2, 2], F2, [1, 2, 3])
    h_one, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [2], 3 * np.ones(20), [1, 2, 3])
    np.testing.assert_array_equal(h_dup, h_one)
    h_b, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [1, 3], np.ones(20), [1, 2, 3])
    h_e, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, [1, 3], np.ones((2, 20)), [1, 2, 3])
    np.testing.assert_array_equal(h_b, h_e)


def test_c_march_matches_literal_oracle():
    mod, matls = coupled_block(2e-3, 1e-3, 0.75e-3, 0.25e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    from conftest import PKG
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n_t = 250
    t = np.arange(n_t) * dt
    F = PKG.mesh.hann_toneburst(t, 5e6, 3)
    fi = [int(gl[PKG.mesh.fn_find_node_at_point(mod["nds"], [1e-3, 1.5e-3]) - 1, 3])]
    hi = np.arange(1, K.shape[0] + 1, 5)
    for mode in MODES:
        h1, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, solver_mode=mode)
        h2, _ = om.OracleMarch(K, C, M, dt, mode).run(n_t, np.array(fi) - 1, F, hi - 1)
        assert rel_l2(h2, h1) < 1e-12


# ---------------------------------------------------------------------------
# elements


def _random_tri(rng):
    while True:
        xy = rng.random((3, 2)) * 1e-3
        J = (xy[1, 0] - xy[0, 0]) * (xy[2, 1] - xy[0, 1]) - (xy[2, 0] - xy[0, 0]) * (xy[1, 1] - xy[0, 1])
        if J > 1e-8:
            return xy


def test_cpe3_rigid_modes_symmetry_patch():
    rng = np.random.default_rng(1)
    D = steel()["D"]
    for _ in range(5):
        xy = _random_tri(rng)
        K, _, M, ln, ld = oe.cpe3(xy, D, 8000.0)
        scale = np.abs(K).max()
        assert np.abs(K - K.T).max() < 1e-12 * scale
        tx = np.tile([1, 0, 0], 3)
        ty = np.tile([0, 1, 0], 3)
        tz = np.tile([0, 0, 1], 3)
        rot = np.concatenate([[-y, x, 0] for x, y in xy])
        for r in (tx, ty, tz, rot):
            assert np.abs(K @ r).max() < 1e-10 * scale * 1e-3
        # patch: linear field u = G x gives nodal forces = integral of B^T sigma
        G = rng.standard_normal((2, 2)) * 1e-3
        u = np.concatenate([[*(G @ p), 0] for p in xy])
        eps = np.array([G[0, 0], G[1, 1], 0, 0, 0, G[0, 1] + G[1, 0]])
        sig = D @ eps
        J = (xy[1, 0] - xy[0, 0]) * (xy[2, 1] - xy[0, 1]) - (xy[2, 0] - xy[0, 0]) * (xy[1, 1] - xy[0, 1])
        # traction check: f_i = (J/2) * B_i^T sigma
        b = [xy[1, 1] - xy[2, 1], xy[2, 1] - xy[0, 1], xy[0, 1] - xy[1, 1]]
        c = [xy[2, 0] - xy[1, 0], xy[0, 0] - xy[2, 0], xy[1, 0] - xy[0, 0]]
        f = np.concatenate([[0.5 * (b[i] * sig[0] + c[i] * sig[5]),
                             0.5 * (c[i] * sig[1] + b[i] * sig[5]), 0.0] for i in range(3)])
        assert np.abs(K @ u - f).max() < 1e-9 * np.abs(f).max()
        assert np.allclose(np.diag(M), J * 8000.0 / 6)


def test_cpe3_quadrature_cross_check():
    """Ke against a 3-point quadrature of B^T D B with shape-function gradients
    obtained by solving for the barycentric coordinates (independent path)."""
    rng = np.random.default_rng(7)
    D = steel()["D"]
    xy = _random_tri(rng)
    K, *_ = oe.cpe3(xy, D, 1.0)
    P = np.hstack([np.ones((3, 1)), xy])          # N_i(x, y) = a_i + b_i x + c_i y
    coef = np.linalg.inv(P)                        # columns: coefficients of N_i
    gx, gy = coef[1], coef[2]
    area = 0.5 * np.linalg.det(P)
    pts = [(2 / 3, 1 / 6, 1 / 6), (1 / 6, 2 / 3, 1 / 6), (1 / 6, 1 / 6, 2 / 3)]
    Kq = np.zeros((9, 9))
    for _ in pts:
        B = np.zeros((6, 9))
        for i in range(3):
            B[0, 3 * i] = gx[i]
            B[1, 3 * i + 1] = gy[i]
            B[3, 3 * i + 2] = gy[i]
            B[4, 3 * i + 2] = gx[i]
            B[5, 3 * i] = gy[i]
            B[5, 3 * i + 1] = gx[i]
        Kq += B.T @ D @ B * area / 3
    assert np.abs(K - Kq).max() < 1e-10 * np.abs(K).max()


def test_cpe3_clockwise_flips_sign():
    rng = np.random.default_rng(3)
    xy = _random_tri(rng)
    D = steel()["D"]
    K1, _, M1, *_ = oe.cpe3(xy, D, 8000.0)
    K2, _, M2, *_ = oe.cpe3(xy[[0, 2, 1]], D, 8000.0)
    perm = np.array([0, 1, 2, 6, 7, 8, 3, 4, 5])
    assert np.allclose(K2, -K1[np.ix_(perm, perm)], rtol=1e-12, atol=1e-3)
    assert np.allclose(np.diag(M2), -np.diag(M1))


def test_ac2d3_null_space_and_mass():
    rng = np.random.default_rng(4)
    xy = _random_tri(rng)
    K, _, M, _, ld = oe.ac2d3(xy, 2.19e9, 1000.0)
    assert ld == [4, 4, 4]
    assert np.abs(K @ np.ones(3)).max() < 1e-12 * np.abs(K).max()
    assert np.all(np.diag(K) < 0)  # -(1/rho) Laplacian
    J = (xy[1, 0] - xy[0, 0]) * (xy[2, 1] - xy[0, 1]) - (xy[2, 0] - xy[0, 0]) * (xy[1, 1] - xy[0, 1])
    assert np.allclose(np.diag(M), -J / (6 * 2.19e9))


def test_asi2d2_normal_translation_and_flip():
    xy = np.array([[0.1e-3, 0.2e-3], [0.5e-3, 0.45e-3]])
    _, C, *_ = oe.asi2d2(xy)
    assert np.abs(C - C.T).max() == 0
    L = np.linalg.norm(xy[1] - xy[0])
    nrm = np.array([xy[1, 1] - xy[0, 1], xy[0, 0] - xy[1, 0]]) / L
    u = np.array([nrm[0], nrm[1], 0, 0, nrm[0], nrm[1], 0, 0])
    r = C @ u
    assert np.allclose(r[[3, 7]], L / 2)
    _, C2, *_ = oe.asi2d2(xy[::-1])
    assert np.allclose(C2, -C[np.ix_([4, 5, 6, 7, 0, 1, 2, 3], [4, 5, 6, 7, 0, 1, 2, 3])])


def test_c3d4_rigid_modes_volume_patch():
    rng = np.random.default_rng(5)
    D = steel()["D"]
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float) * 1e-3
    X = X + rng.random((4, 3)) * 2e-4
    K, _, M, _, _ = oe.c3d4(X, D, 8000.0)
    V = np.linalg.det(np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6
    assert V > 0
    assert np.allclose(np.diag(M), 8000.0 * V / 4)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-12 * scale
    modes = [np.tile(e, 4) for e in np.eye(3)]
    for w in np.eye(3):
        modes.append(np.concatenate([np.cross(w, p) for p in X]))
    for r in modes:
        assert np.abs(K @ r).max() < 1e-9 * scale * 1e-3
    G = rng.standard_normal((3, 3)) * 1e-3
    u = np.concatenate([G @ p for p in X])
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0], G[0, 1] + G[1, 0]])
    E = 0.5 * u @ K @ u
    assert np.isclose(E, 0.5 * eps @ D @ eps * V, rtol=1e-10)


# ---------------------------------------------------------------------------
# post-processing / stability


def test_ke_field_output_closed_form():
    """One element, uniform velocity v in x: sqrt(KE) = sqrt(mean_nodes(m v^2 / 2))."""
    nds = np.array([[0, 0], [1, 0], [0, 1]], dtype=float) * 1e-3
    mod = {"nds": nds, "els": np.array([[1, 2, 3]]), "el_mat_i": np.array([1]),
           "el_abs_i": np.zeros(1), "el_typ_i": np.array(["CPE3"], dtype=object)}
    K, C, M, gl = oa.build_global_matrices(nds, mod["els"], [1], [0.0], ["CPE3"], [steel()],
                                           {"dof_to_use": [1, 2]})
    v = np.zeros((K.shape[0], 1))
    v[gl[:, 0] - 1, 0] = 2.0
    out = op.field_output_ke(v, mod["els"], M.diagonal(), gl)
    m_node = 1e-6 * 8000 / 6
    assert np.isclose(out[0, 0], np.sqrt(m_node * 4 / 2))


def test_amplification_matrix_spectral_radius_matches_dt_limit():
    mod, matls = solid_plate(1e-3, 0.5e-3, 0.25e-3)
    K, C, M, _ = assemble(mod, matls, {"dof_to_use": [1, 2]})
    lam = np.max(np.linalg.eigvalsh((np.diag(1 / np.sqrt(M.diagonal())) @ K.toarray()
                                     @ np.diag(1 / np.sqrt(M.diagonal())))))
    dt_crit = 2 / np.sqrt(lam)
    for mode in MODES:
        # rigid-body modes give a defective eigenvalue 1 (Jordan block): 1 + O(sqrt(eps))
        assert op.max_amplification(M, C, K, 0.95 * dt_crit, mode) <= 1 + 1e-6
        assert op.max_amplification(M, C, K, 1.05 * dt_crit, mode) > 1 + 1e-3


@pytest.mark.parametrize("mode", ["vel at curent time step", "vel at last half time step"])
def test_abd_diagonal_x_equals_generic_formation(mode):
    """The row-block formation of A, B, D for a diagonal C (the CPU baseline's
    literal march on large models) gives the generic sparse formation's
    matrices (v6:130-139) to rounding, on an absorbing plate, and the same
    histories through the literal loop; t_end stops the loop after that
    step."""
    mod, matls = solid_plate(2e-3, 1e-3, 0.1e-3)
    inner = np.array([[0.3e-3, 0.3e-3], [1.7e-3, 0.3e-3], [1.7e-3, 5e-3], [0.3e-3, 5e-3]])
    mod = PKG.mesh.fn_add_absorbing_layer(mod, inner, 0.3e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2], "damping_power_law": 2})
    assert C.diagonal().max() > 0
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    ref = ov6.abd_matrices(K, C, M, dt, mode)
    got = ov6.abd_diagonal_x(K, C, M, dt, mode, chunk=97)
    for a, b in zip(got, ref):
        assert abs(a - b).max() <= 1e-15 * abs(b).max()
    t = np.arange(200) * dt
    F = PKG.mesh.hann_toneburst(t, 5e6, 3)
    fi, hi = [gl[50, 1]], np.arange(1, K.shape[0] + 1, 7)
    h1, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, solver_mode=mode, abd=ref)
    h2, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, solver_mode=mode, abd=got,
                                                 t_end=120)
    assert np.linalg.norm(h1[:, :120] - h2[:, :120]) <= 1e-12 * np.linalg.norm(h1[:, :120])
    assert not h2[:, 120:].any() and h1[:, 120:].any()
    Cx = C.tolil()
    Cx[0, 1] = 1.0
    with pytest.raises(ValueError):
        ov6.abd_diagonal_x(K, Cx.tocsr(), M, dt, mode)

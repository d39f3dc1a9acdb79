"""Pins of the oracle's subdomain operations (oracle/subdoms.py) by what the
mathematics fixes: direct convolution (numpy.convolve), the circular sum that
FFTs of length 2n compute, and the v6 recurrence read backwards (forces from
a marched displacement history give back the applied forcing)."""
import numpy as np
import pytest

from conftest import PKG, assemble, solid_plate
from oracle import subdoms as osd
from oracle import v6 as ov6


def test_convolve_is_truncated_linear_convolution():
    rng = np.random.default_rng(0)
    f = rng.standard_normal((3, 37))
    g = rng.standard_normal(37)
    h = osd.convolve(f, g[None, :], 2)
    for s in range(3):
        np.testing.assert_allclose(h[s], np.convolve(f[s], g)[:37], rtol=0, atol=1e-12)
    # along the columns, with one g per signal
    G = rng.standard_normal((37, 3))
    h1 = osd.convolve(f.T, G, 1)
    for s in range(3):
        np.testing.assert_allclose(h1[:, s], np.convolve(f[s], G[:, s])[:37], rtol=0, atol=1e-12)
    # an impulse is the identity
    d = np.zeros(37)
    d[0] = 1.0
    np.testing.assert_allclose(osd.convolve(f, d[None, :], 2), f, rtol=0, atol=1e-14)


def test_convolve_long_g_is_circular_of_length_2n():
    rng = np.random.default_rng(1)
    n, P = 11, 22
    f = rng.standard_normal((1, n))
    g = rng.standard_normal(40)            # truncated to P samples by fft(g, P)
    h = osd.convolve(f, g[None, :], 2)[0]
    gp = g[:P]
    brute = np.array([sum(f[0, k] * gp[(t - k) % P] for k in range(n)) for t in range(n)])
    np.testing.assert_allclose(h, brute, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", ["vel at curent time step", "vel at last half time step"])
def test_forces_from_marched_history_give_back_the_forcing(mode):
    mod, matls = solid_plate(1e-3, 0.6e-3, 0.1e-3)
    inner = np.array([[0.2e-3, 0.2e-3], [0.8e-3, 0.2e-3], [0.8e-3, 5e-3], [0.2e-3, 5e-3]])
    mod = PKG.mesh.fn_add_absorbing_layer(mod, inner, 0.2e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2], "damping_power_law": 2})
    n = K.shape[0]
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    t = np.arange(120) * dt
    F = PKG.mesh.hann_toneburst(t, 8e6, 2)
    fi = np.array([gl[PKG.mesh.fn_find_node_at_point(mod["nds"], [0.5e-3, 0.3e-3]) - 1, 1]])
    hist, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, np.arange(1, n + 1),
                                                   np.inf, "KE", mode)
    lyrs = np.full(n, 2)                     # every DOF forced, none frozen ('in')
    force, fs, fin, fout = osd.disps_to_forces(K, C, M, dt, hist, lyrs, "in", mode)
    assert fs.all() and fin.all() and not fout.any()
    expect = np.zeros((n, t.size))
    expect[fi[0] - 1] = F
    assert np.abs(force - expect).max() <= 1e-9 * np.abs(F).max()
    # a plausible slip (velocity or stiffness term one step off) breaks it
    force2, *_ = osd.disps_to_forces(K, C, M, dt, np.roll(hist, 1, axis=1), lyrs, "in", mode)
    assert np.abs(force2 - expect).max() > 1e-3 * np.abs(F).max()


def test_forces_freeze_the_named_layers():
    rng = np.random.default_rng(2)
    mod, matls = solid_plate(0.5e-3, 0.3e-3, 0.1e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2]})
    n = K.shape[0]
    disp = rng.standard_normal((n, 9))
    lyrs = rng.integers(0, 5, n)
    for side, frozen in (("in", (3, 4)), ("out", (1, 2))):
        f1, fs, _, _ = osd.disps_to_forces(K, C, M, 1e-8, disp, lyrs, side,
                                           "vel at curent time step")
        d0 = disp.copy()
        d0[np.isin(lyrs, frozen)] = 0.0
        f2, _, _, _ = osd.disps_to_forces(K, C, M, 1e-8, d0, np.full(n, 2), "in",
                                          "vel at curent time step")
        np.testing.assert_array_equal(np.nonzero(fs)[0], np.nonzero(np.isin(lyrs, (2, 3)))[0])
        np.testing.assert_allclose(f1, f2[fs], rtol=0, atol=1e-9 * np.abs(f2).max())


def test_parse_to_fmc_sums_or_weights_receiver_rows():
    steps = [{"mon": {"nds": np.array([1, 1, 2, 2, 5, 5])}} for _ in range(2)]
    rows = np.arange(6 * 4, dtype=float).reshape(6, 4)
    res = [{"dsps": rows}, {"dsps": 2 * rows}]
    trans = [{"nds": [1, 2]}, {"nds": [5], "wt": [0.5, -1.0]}]
    tx, rx = np.array([1, 2, 1, 2]), np.array([1, 1, 2, 2])
    td = osd.parse_to_fmc(tx, rx, 4, steps, res, trans, None)
    np.testing.assert_array_equal(td[:, 0], rows[:4].sum(axis=0))
    np.testing.assert_array_equal(td[:, 1], 2 * rows[:4].sum(axis=0))
    np.testing.assert_array_equal(td[:, 2], 0.5 * rows[4] - rows[5])
    np.testing.assert_array_equal(td[:, 3], 2 * (0.5 * rows[4] - rows[5]))


def _subdomain_case(scatterer):
    m = PKG.mesh
    mod, matls = solid_plate(6e-3, 3e-3, 0.1e-3)
    inner = np.array([[0.5e-3, 0.5e-3], [5.5e-3, 0.5e-3], [5.5e-3, 5e-3], [0.5e-3, 5e-3]])
    mod = m.fn_add_absorbing_layer(mod, inner, 0.5e-3)
    box = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]]) * 0.5e-3 + np.array([3e-3, 1.6e-3])
    dm = PKG.subdoms.fn_create_subdomain(mod, matls, box, 0.5e-3)
    if scatterer:
        blob = m.fn_create_smooth_random_blob(0.5, 3, 64, np.random.default_rng(9)) * 0.3e-3
        dm = m.fn_add_scatterer(dm, matls, blob + np.array([3e-3, 1.6e-3]), 0)
    return mod, matls, box, dm


def test_create_subdomain_structure():
    """fn_create_subdomain.m:42-87 on a structured plate: the subdomain's
    nodes are main-model nodes (main_nd_i), the inner elements keep no
    absorbing index, each boundary layer i >= 2 lies next to layer i - 1, the
    layers grow outward, the absorbing index stays in [0, 1], every inner
    element is kept, and a void scatterer removes elements and nodes while
    main_nd_i and bdry_lyrs follow the nodes."""
    mod, matls, box, dm = _subdomain_case(False)
    nds = np.asarray(mod["nds"])
    np.testing.assert_array_equal(dm["nds"], nds[dm["main_nd_i"] - 1])
    lyrs = dm["bdry_lyrs"]
    counts = [int(np.count_nonzero(lyrs == i)) for i in range(1, 5)]
    assert all(c > 0 for c in counts) and counts == sorted(counts)
    els = dm["els"]
    for i in range(2, 5):
        # every layer-i node shares an element with a layer-(i-1) node
        have = np.isin(els, np.nonzero(lyrs == i - 1)[0] + 1).any(axis=1)
        near = np.unique(els[have])
        assert np.all(np.isin(np.nonzero(lyrs == i)[0] + 1, near))
    a = dm["el_abs_i"]
    assert a.min() >= 0 and a.max() <= 1 and a.max() > 0.9
    assert not a[dm["int_el_i"]].any()
    # the main model's inner elements are all in the subdomain
    n_inner = dm["main_int_el_i"].size
    assert n_inner == int(np.count_nonzero(dm["int_el_i"])) > 0
    # layer-1 nodes are nodes of the inner elements
    inner_nodes = np.unique(els[dm["int_el_i"]])
    assert np.all(np.isin(np.nonzero(lyrs == 1)[0] + 1, inner_nodes))
    _, _, _, dv = _subdomain_case(True)
    assert dv["els"].shape[0] < els.shape[0] and dv["nds"].shape[0] < dm["nds"].shape[0]
    np.testing.assert_array_equal(dv["nds"], nds[dv["main_nd_i"] - 1])
    kept = np.isin(dm["main_nd_i"], dv["main_nd_i"])
    np.testing.assert_array_equal(dv["bdry_lyrs"], lyrs[kept])

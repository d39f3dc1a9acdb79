"""Subdomain coupling operations on the GPU against the oracle
(code/internal/fn_convolve.m, code/fn_convert_disps_to_forces_v2.m,
subdoms/fn_run_main_model.m)."""
import numpy as np
import pytest

from conftest import PKG, assemble, coupled_block, rel_l2, solid_plate
from oracle import subdoms as osd
from oracle import v6 as ov6

pytestmark = pytest.mark.gpu
MODES = ["vel at curent time step", "vel at last half time step"]


def _close(a, b, rtol):
    assert a.shape == b.shape
    assert np.abs(a - b).max() <= rtol * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("shape,dim,per,n_g", [
    ((5, 257), 2, False, 257), ((300, 64), 1, False, 300), ((7, 100), 2, True, 100),
    ((3, 41), 2, False, 200), ((1, 1), 2, False, 1), ((60, 200_000), 2, False, 200_000)])
def test_convolve_matches_oracle(shape, dim, per, n_g):
    print("case: convolve", shape, dim, per, n_g, flush=True)
    rng = np.random.default_rng(sum(shape) + n_g)
    f = rng.standard_normal(shape)
    n_sig = shape[0] if dim == 2 else shape[1]
    if per:
        g = rng.standard_normal((n_sig, n_g)) if dim == 2 else rng.standard_normal((n_g, n_sig))
        go = g
    else:
        g = rng.standard_normal(n_g)
        go = g[None, :] if dim == 2 else g[:, None]
    h = PKG.subdoms.fn_convolve(f, g, dim)
    _close(h, osd.convolve(f, go, dim), 1e-12)


def test_convolve_vector_and_impulse():
    rng = np.random.default_rng(4)
    f = rng.standard_normal(333)
    d = np.zeros(333)
    d[0] = 1.0
    np.testing.assert_allclose(PKG.subdoms.fn_convolve(f, d, 1), f, rtol=0, atol=1e-13)
    g = rng.standard_normal(333)
    _close(PKG.subdoms.fn_convolve(f, g, 2), np.convolve(f, g)[:333], 1e-12)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("side", ["in", "out"])
def test_disps_to_forces_matches_oracle(mode, side):
    """Coupled model (C with ASI2D2 couplings and absorbing damping), random
    histories and layers."""
    print("case: disps to forces", mode, side, flush=True)
    mod, matls = coupled_block(1.5e-3, 0.75e-3, 0.5e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    n = K.shape[0]
    rng = np.random.default_rng(7)
    disp = rng.standard_normal((n, 50))
    lyrs = rng.integers(0, 5, n)
    dt = 1e-8
    f, fs, fin, fout = PKG.subdoms.fn_convert_disps_to_forces_v2(K, C, M, dt, disp, lyrs, side,
                                                                 mode)
    fo, fso, fino, fouto = osd.disps_to_forces(K, C, M, dt, disp, lyrs, side, mode)
    for a, b in ((fs, fso), (fin, fino), (fout, fouto)):
        np.testing.assert_array_equal(a, b)
    _close(f, fo, 1e-13)


@pytest.mark.parametrize("mode", MODES)
def test_forces_from_gpu_march_give_back_the_forcing(mode):
    print("case: forces from the GPU march", mode, flush=True)
    mod, matls = solid_plate(1e-3, 0.6e-3, 0.1e-3)
    inner = np.array([[0.2e-3, 0.2e-3], [0.8e-3, 0.2e-3], [0.8e-3, 5e-3], [0.2e-3, 5e-3]])
    mod = PKG.mesh.fn_add_absorbing_layer(mod, inner, 0.2e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2], "damping_power_law": 2})
    n = K.shape[0]
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    t = np.arange(150) * dt
    F = PKG.mesh.hann_toneburst(t, 8e6, 2)
    fi = np.array([gl[PKG.mesh.fn_find_node_at_point(mod["nds"], [0.5e-3, 0.3e-3]) - 1, 1]])
    hist, _, _, _ = PKG.fn_explicit_dynamic_solver_v6(K, C, M, t, fi, F, [], [],
                                                      np.arange(1, n + 1), np.inf, 1, "KE", mode)
    force, *_ = PKG.subdoms.fn_convert_disps_to_forces_v2(K, C, M, dt, hist, np.full(n, 2), "in",
                                                          mode)
    expect = np.zeros((n, t.size))
    expect[fi[0] - 1] = F
    assert np.abs(force - expect).max() <= 1e-9 * np.abs(F).max()


def test_pristine_fmc_matches_oracle_per_transmitter_runs():
    """fn_run_main_model (impulse response, batched march, GPU parse) against
    one oracle v6 march per transmitter and the oracle's parse."""
    print("case: pristine FMC", flush=True)
    m = PKG.mesh
    mod, matls = solid_plate(4e-3, 2e-3, 0.1e-3)
    inner = np.array([[0.5e-3, 0.5e-3], [3.5e-3, 0.5e-3], [3.5e-3, 5e-3], [0.5e-3, 5e-3]])
    mod = m.fn_add_absorbing_layer(mod, inner, 0.5e-3)
    nds = mod["nds"]
    top = np.nonzero(np.isclose(nds[:, 1], 2e-3))[0] + 1
    trans = []
    for k, xc in enumerate((1.4e-3, 1.8e-3, 2.2e-3, 2.6e-3)):
        sel = top[np.abs(nds[top - 1, 0] - xc) <= 0.1e-3 + 1e-12]
        tr = {"nds": sel, "dfs": np.full(sel.size, 2)}
        if k == 3:
            tr["wt"] = np.linspace(0.5, 1.5, 2 * sel.size)  # one weight per monitored row
        trans.append(tr)
    K0, C0, M0, gl0 = assemble(mod, matls, {"dof_to_use": [1, 2]})
    dt = m.stable_time_step(K0, M0.diagonal())
    opts = {"centre_freq": 5e6, "number_of_cycles": 3, "time_step": dt, "time_pts": 400,
            "dof_to_use": [1, 2], "damping_power_law": 2}
    main = PKG.subdoms.fn_run_main_model({"mod": mod, "matls": matls, "trans": trans}, opts)
    res = main["res"]
    K, C, M, gl = res["mats"]["K"], res["mats"]["C"], res["mats"]["M"], res["mats"]["gl_lookup"]
    t, sig = main["inp"]["time"], main["inp"]["sig"]
    imp = np.zeros(t.size)
    imp[0] = 1.0
    mon_nds, mon_dfs, hi = res["mon_nds"], res["mon_dfs"], res["dsp_gi"]
    steps, ores = [], []
    for e, tr in enumerate(trans):
        fi = gl[np.asarray(tr["nds"]) - 1, 1]
        h, _, _, _ = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, imp, hi)
        assert rel_l2(res["trans"][e]["dsps"], h) < 1e-10
        steps.append({"mon": {"nds": mon_nds, "dfs": mon_dfs}})
        ores.append({"dsps": h})
    fmc = res["fmc"]
    td = osd.parse_to_fmc(fmc["tx"], fmc["rx"], t.size, steps, ores, trans, sig)
    assert fmc["time_data"].shape == (t.size, 16)
    assert rel_l2(fmc["time_data"], td) < 1e-10
    assert np.array_equal(fmc["tx"], np.repeat(np.arange(1, 5), 4))
    assert np.array_equal(fmc["rx"], np.tile(np.arange(1, 5), 4))


def _subdomain_main(scatterer):
    from test_subdoms_oracle import _subdomain_case
    m = PKG.mesh
    mod, matls, box, dm = _subdomain_case(scatterer)
    nds = mod["nds"]
    top = np.nonzero(np.isclose(nds[:, 1], 3e-3))[0] + 1
    trans = []
    for xc in (2.0e-3, 2.7e-3, 3.4e-3, 4.1e-3):
        sel = top[np.abs(nds[top - 1, 0] - xc) <= 0.1e-3 + 1e-12]
        trans.append({"nds": sel, "dfs": np.full(sel.size, 2)})
    K0, C0, M0, gl0 = assemble(mod, matls, {"dof_to_use": [1, 2]})
    dt = m.stable_time_step(K0, M0.diagonal())
    opts = {"centre_freq": 4e6, "number_of_cycles": 3, "time_step": dt, "time_pts": 500,
            "dof_to_use": [1, 2], "damping_power_law": 2}
    main = PKG.subdoms.fn_run_main_model({"mod": mod, "matls": matls, "trans": trans,
                                          "doms": [{"mod": dm}]}, opts)
    return main, opts


def test_subdomain_model_matches_oracle_steps():
    """fn_run_subdomain_model (all transmitters in one batched march of the
    subdomain, forces and convolutions on the GPU) against the oracle's
    line-by-line restatement (one v6 march per transmitter on the oracle's
    assembly, numpy FFT convolution, scipy forces) on a plate with a void
    scatterer in the subdomain: the FMC with the scattered field added agrees
    to 1e-10, and the scattered part is a real signal."""
    print("case: subdomain model with a void", flush=True)
    from oracle import assembly as oa
    main, opts = _subdomain_main(True)
    sub_opts = {"dof_to_use": [1, 2], "damping_power_law": 2}
    main2 = PKG.subdoms.fn_run_subdomain_model(main, sub_opts)
    got = main2["doms"][0]["res"]["fmc"]["time_data"]
    pristine = main["res"]["fmc"]["time_data"]
    o = {"tx_trans": [1, 2, 3, 4], "rx_trans": [1, 2, 3, 4], "doms_to_run": [1],
         "solver_mode": "vel at curent time step", "dof_to_use": [1, 2], "damping_power_law": 2,
         "max_damping": 3.1415e7, "max_stiffness_reduction": 0.01}

    def build(dm, matls, opts_):
        K, C, M, gl = oa.build_global_matrices(dm["nds"], dm["els"], dm["el_mat_i"], dm["el_abs_i"],
                                               list(dm["el_typ_i"]), matls, opts_)
        return K, C, M, gl

    def march(K, C, M, t, fgi, frcs, hgi):
        return ov6.explicit_dynamic_solver_v6(K, C, M, t, fgi, frcs, hgi)[0]

    ref = osd.run_subdomain_model(main, o, build, march)[0]
    scat = got - pristine
    assert np.linalg.norm(scat) > 1e-3 * np.linalg.norm(pristine)
    assert rel_l2(got, ref) < 1e-10
    assert rel_l2(scat, ref - pristine) < 1e-8
    # a subdomain without a scatterer scatters (almost) nothing: the injected
    # incident field stays inside the layers
    main_p, _ = _subdomain_main(False)
    m3 = PKG.subdoms.fn_run_subdomain_model(main_p, sub_opts)
    s0 = m3["doms"][0]["res"]["fmc"]["time_data"] - main_p["res"]["fmc"]["time_data"]
    print(f"scattered / pristine: void {np.linalg.norm(scat) / np.linalg.norm(pristine):.3e}, "
          f"none {np.linalg.norm(s0) / np.linalg.norm(pristine):.3e}")
    assert np.linalg.norm(s0) < 0.05 * np.linalg.norm(scat)

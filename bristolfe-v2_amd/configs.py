"""Seeded synthetic models of the benchmark workloads.

Each builder returns a dict with the model (``mod``, ``matls``,
``fe_options``), the assembled ``K, C, M, gl_lookup``, the time axis, the
forcing (1-based global indices + signals) and the monitored DOFs.

  plate_2d      2D plane-strain steel plate, CPE3 (DOF 1, 2), one toneburst
                force at the top-centre node (workload of 20 mm x 10 mm at
                0.1 mm: 20,301 nodes, 40,000 elements, 40,602 DOF).
  coupled_2d    aluminium block under a water layer with ASI2D2 interface
                elements, 5 mm absorbing layers (quadratic ramp) on the left,
                right and bottom, a seeded smooth-blob void and a toneburst
                pressure source in the water (60 mm x 30 mm + 20 mm water at
                0.025 mm: 7,690,403 DOF before the void).
  array_plate   steel plate with a linear array of surface sources (one
                source per element, full-matrix capture).

Time step: 0.9 * 2 / omega_max, omega_max^2 the largest eigenvalue of
M^-1 K, estimated on a patch of the same element size and materials (the
largest eigenvalue is a local, element-scale property; damping and the
absorbing stiffness reduction only lower it).
"""
from __future__ import annotations

import numpy as np

from . import mesh as _m
from .assembly import fn_build_global_matrices_v5


def steel():
    return {"rho": 8000.0, "D": _m.fn_isotropic_stiffness_matrix(210e9, 0.3), "el_typ": "CPE3",
            "name": "Steel"}


def aluminium():
    return {"rho": 2700.0, "D": _m.fn_isotropic_stiffness_matrix(70e9, 0.33), "el_typ": "CPE3",
            "name": "Aluminium"}


def water():
    return {"rho": 1000.0, "D": 1480.0 ** 2 * 1000.0, "el_typ": "AC2D3", "name": "Water"}


def _rect(x0, y0, x1, y1):
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], dtype=np.float64)


def _assemble(mod, matls, opts):
    return fn_build_global_matrices_v5(mod["nds"], mod["els"], mod["el_mat_i"], mod["el_abs_i"],
                                       mod["el_typ_i"], matls, opts)


def _typ(mod, matls):
    mod["el_typ_i"] = np.asarray([matls[i - 1]["el_typ"] for i in mod["el_mat_i"]], dtype=object)
    return mod


def patch_time_step(matls, el_size, dof_to_use, safety=0.9):
    """Stable dt from a 24 x 24 element patch of every material."""
    dts = []
    for k, mat in enumerate(matls):
        mod = _m.fn_rectangular_structured_mesh(_rect(0, 0, 24 * el_size, 24 * el_size), el_size)
        mod["el_mat_i"][:] = 1
        mod["el_typ_i"] = np.array([mat["el_typ"]] * len(mod["el_mat_i"]), dtype=object)
        K, C, M, _ = _assemble(mod, [mat], {"dof_to_use": dof_to_use})
        dts.append(_m.stable_time_step(K, M.diagonal(), safety=safety, iters=2000, tol=1e-12))
    return min(dts)


def plate_2d(width=20e-3, height=10e-3, el_size=0.1e-3, n_t=3000, centre_freq=2e6, cycles=5,
             field_every=100):
    matls = [steel()]
    mod = _typ(_m.fn_rectangular_structured_mesh(_rect(0, 0, width, height), el_size), matls)
    opts = {"dof_to_use": [1, 2], "field_output_every_n_frames": field_every}
    K, C, M, gl = _assemble(mod, matls, opts)
    dt = _m.stable_time_step(K, M.diagonal(), iters=3000, tol=1e-12)
    t = np.arange(n_t) * dt
    sig = _m.hann_toneburst(t, centre_freq, cycles)
    nd = _m.fn_find_node_at_point(mod["nds"], [width / 2, height])
    gi = np.array([gl[nd - 1, 1]])
    return dict(mod=mod, matls=matls, fe_options=opts, K=K, C=C, M=M, gl_lookup=gl, time=t, dt=dt,
                frc_gi=gi, frcs=sig[None, :], hist_gi=gi.copy(), src_node=nd,
                name=f"plate_2d {width*1e3:g}x{height*1e3:g} mm h={el_size*1e3:g} mm")


def coupled_2d(width=60e-3, solid_h=30e-3, water_h=20e-3, el_size=0.025e-3, layer=5e-3,
               n_t=8000, centre_freq=5e6, cycles=5, n_mon=64, void_radius=1.5e-3, seed=1234,
               dt=None):
    rng = np.random.default_rng(seed)
    matls = [aluminium(), water()]
    mod = _m.fn_rectangular_structured_mesh(_rect(0, 0, width, solid_h + water_h), el_size)
    mod = _m.fn_set_els_inside_bdry_to_mat(mod, _rect(-1, solid_h, width + 1, solid_h + water_h + 1), 2)
    mod = _typ(mod, matls)
    if void_radius > 0:
        blob = _m.fn_create_smooth_random_blob(0.5, 3, 96, rng) * void_radius
        cx = width * (0.35 + 0.3 * rng.random())
        cy = solid_h * (0.35 + 0.3 * rng.random())
        mod = _m.remove_elements_in_region(mod, blob + np.array([cx, cy]))
    inner = _rect(layer, layer, width - layer, 10 * (solid_h + water_h))
    mod = _m.fn_add_absorbing_layer(mod, inner, layer)
    mod = _m.fn_add_fluid_solid_interface_els(mod, matls)
    opts = {"dof_to_use": [1, 2, 4], "damping_power_law": 2}
    K, C, M, gl = _assemble(mod, matls, opts)
    if dt is None:
        dt = patch_time_step(matls, el_size, [1, 2, 4])
    t = np.arange(n_t) * dt
    sig = _m.hann_toneburst(t, centre_freq, cycles)
    sx = width * (0.4 + 0.2 * rng.random())
    sy = solid_h + water_h * (0.4 + 0.3 * rng.random())
    nd = _m.fn_find_node_at_point(mod["nds"], [sx, sy])
    frc = np.array([gl[nd - 1, 3]])
    xs = width / 2 + (np.arange(n_mon) - (n_mon - 1) / 2) * (0.5e-3 if width >= 40e-3 else width / (2 * n_mon))
    ym = solid_h - 1e-3 if solid_h > 2e-3 else solid_h / 2
    kk = _m.find_nodes_at_points(mod["nds"], np.stack([xs, np.full(xs.size, ym)], axis=1))
    mon = gl[kk - 1, 1]
    mon = mon[mon > 0]
    return dict(mod=mod, matls=matls, fe_options=opts, K=K, C=C, M=M, gl_lookup=gl, time=t, dt=dt,
                frc_gi=frc, frcs=sig[None, :], hist_gi=mon, src_node=nd,
                name=(f"coupled_2d Al {width*1e3:g}x{solid_h*1e3:g} mm + water {water_h*1e3:g} mm "
                      f"h={el_size*1e3:g} mm"))


def array_plate(width=20e-3, height=10e-3, el_size=0.1e-3, n_el=64, pitch=0.3e-3, n_t=3000,
                centre_freq=2e6, cycles=5, dt=None):
    """Linear array on the top surface: element e = the surface node nearest
    its centre, DOF 2. Returns per-source forcing (frc_gi, frc_src)."""
    matls = [steel()]
    mod = _typ(_m.fn_rectangular_structured_mesh(_rect(0, 0, width, height), el_size), matls)
    opts = {"dof_to_use": [1, 2]}
    K, C, M, gl = _assemble(mod, matls, opts)
    if dt is None:
        dt = _m.stable_time_step(K, M.diagonal(), iters=3000, tol=1e-12)
    t = np.arange(n_t) * dt
    sig = _m.hann_toneburst(t, centre_freq, cycles)
    xs = width / 2 + (np.arange(n_el) - (n_el - 1) / 2) * pitch
    nodes = _m.find_nodes_at_points(mod["nds"], np.stack([xs, np.full(xs.size, height)], axis=1))
    gi = gl[nodes - 1, 1]
    return dict(mod=mod, matls=matls, fe_options=opts, K=K, C=C, M=M, gl_lookup=gl, time=t, dt=dt,
                el_gi=gi, frcs=sig[None, :], nodes=nodes,
                name=f"array_plate {n_el} el pitch {pitch*1e3:g} mm")

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
  cube_3d       steel cube of linear tetrahedra (C3D4, 6 per cubic cell),
                absorbing layers on the four side faces, a toneburst normal
                force on a 3 x 3-node patch at the top-face centre and 16
                top-face monitors (20 mm at 0.1 mm: 8,120,601 nodes,
                48,000,000 tets, 24,361,803 DOF).

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


def patch_time_step_tets(mat, el_size, safety=0.9, cells=8):
    """Stable dt from a cells^3 patch of the structured tet mesh."""
    mod = _m.cubic_structured_tet_mesh(cells * el_size, el_size)
    mod["el_typ_i"] = np.array(["C3D4"] * len(mod["el_mat_i"]), dtype=object)
    K, C, M, _ = _assemble(mod, [mat], {"dof_to_use": []})
    return _m.stable_time_step(K, M.diagonal(), safety=safety, iters=2000, tol=1e-12)


def steel_3d():
    return dict(steel(), el_typ="C3D4")


def cube_3d(size=20e-3, el_size=0.1e-3, layer=2e-3, n_t=2000, centre_freq=2e6, cycles=3,
            n_mon=16, mon_pitch=1.5e-3, dt=None, assembly_backend="auto", vertical=2):
    """Steel cube of C3D4 tets (cubic_structured_tet_mesh: nodes numbered x
    fastest, then y, then z) with absorbing layers on the four faces parallel
    to the vertical axis, a normal toneburst force on a 3 x 3-node patch at
    the centre of the top face (the face at the top of the vertical axis) and
    n_mon monitors on a square grid around it on that face. vertical = 2 (the
    default): layers on the x and y faces, top face z = size. vertical = 0
    makes x, the fastest node index, the vertical axis (layers on the y and z
    faces): the same physical model, rotated, whose node runs along x then
    end on the two free x faces instead of in the layers."""
    matls = [steel_3d()]
    mod = _m.cubic_structured_tet_mesh(size, el_size)
    n_el = mod["els"].shape[0]
    mod["el_typ_i"] = np.full(n_el, "C3D4", dtype=object)
    if vertical not in (0, 2):
        raise ValueError("vertical axis must be 0 (x) or 2 (z)")
    lo = [layer] * 3
    hi = [size - layer] * 3
    lo[vertical], hi[vertical] = -np.inf, np.inf
    mod = _m.add_absorbing_layer_box(mod, tuple(lo), tuple(hi), layer)
    opts = {"dof_to_use": [], "damping_power_law": 2, "assembly_backend": assembly_backend}
    K, C, M, gl = _assemble(mod, matls, opts)
    if dt is None:
        dt = patch_time_step_tets(matls[0], el_size)
    t = np.arange(n_t) * dt
    sig = _m.hann_toneburst(t, centre_freq, cycles)
    nds = mod["nds"]
    nc = int(round(size / el_size))
    nx = nc + 1

    def node(i, j, k):  # grid indices along x, y, z -> 0-based node
        return i + nx * (j + nx * k)

    # in-face axes (a, b) of the top face, the vertical index at its top
    a_ax, b_ax = [d for d in range(3) if d != vertical]

    def face_node(ia, ib):
        g = [0, 0, 0]
        g[vertical], g[a_ax], g[b_ax] = nc, ia, ib
        return node(*g)

    ic = nc // 2
    patch = np.array([face_node(ic + a, ic + b) for b in (-1, 0, 1) for a in (-1, 0, 1)])
    assert np.allclose(nds[patch, vertical], size)
    frc = gl[patch, vertical]
    k = int(round(np.sqrt(n_mon)))
    off = (np.arange(k) - (k - 1) / 2) * mon_pitch
    A, B = np.meshgrid(size / 2 + off, size / 2 + off)
    ii = np.rint(A.ravel() / el_size).astype(np.int64)
    jj = np.rint(B.ravel() / el_size).astype(np.int64)
    mon_nodes = np.array([face_node(p, q) for p, q in zip(ii, jj)], dtype=np.int64)
    mon = gl[mon_nodes, vertical]
    return dict(mod=mod, matls=matls, fe_options=opts, K=K, C=C, M=M, gl_lookup=gl, time=t, dt=dt,
                frc_gi=frc, frcs=sig[None, :], hist_gi=mon, src_node=patch + 1,
                name=f"cube_3d steel {size*1e3:g} mm C3D4 h={el_size*1e3:g} mm")


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


def adjoint_plate(width=30e-3, height=15e-3, el_size=0.05e-3, n_el=16, pitch=1e-3, centre_freq=3e6,
                  cycles=5, n_t=5000, ckpt_every=250, incl_radius=2.0e-3, seed=77):
    """Steel plate with a 16-element surface array for gradient-based
    inversion: the parametrised model, the shots (each array element one
    transmitter, DOF 2; all elements receive), and the target parameters of
    a seeded smooth-blob inclusion (stiffness x 0.7, density x 0.9)."""
    from .adjoint import ParametrizedModel, Shots

    rng = np.random.default_rng(seed)
    matls = [steel()]
    mod = _typ(_m.fn_rectangular_structured_mesh(_rect(0, 0, width, height), el_size), matls)
    pm = ParametrizedModel(mod, matls, {"dof_to_use": [1, 2]})
    dt = _m.stable_time_step(pm.K0, pm.M0.diagonal(), iters=3000, tol=1e-12)
    t = np.arange(n_t) * dt
    sig = _m.hann_toneburst(t, centre_freq, cycles)
    xs = width / 2 + (np.arange(n_el) - (n_el - 1) / 2) * pitch
    nodes = _m.find_nodes_at_points(mod["nds"], np.stack([xs, np.full(xs.size, height)], axis=1))
    dofs = pm.gl[nodes - 1, 1] - 1
    shots = Shots([np.array([d]) for d in dofs], sig, dofs, n_t, dt, ckpt_every)
    blob = _m.fn_create_smooth_random_blob(0.5, 3, 64, rng) * incl_radius
    centre = np.array([width * (0.4 + 0.2 * rng.random()), height * (0.35 + 0.2 * rng.random())])
    ctr = _m.fn_calc_element_centres(mod["nds"], mod["els"])[pm.el_index]
    inside = _m.inpolygon(ctr[:, 0], ctr[:, 1], blob[:, 0] + centre[0], blob[:, 1] + centre[1])
    s_true = np.where(inside, 0.7, 1.0)
    rho_true = np.where(inside, 0.9, 1.0) * pm.rho0
    return dict(model=pm, shots=shots, s_true=s_true, rho_true=rho_true, inside=inside, dt=dt,
                name=(f"adjoint_plate steel {width*1e3:g}x{height*1e3:g} mm h={el_size*1e3:g} mm, "
                      f"{n_el} tx/rx, {n_t} steps"))


def _whole(k, el_size):
    """The largest length of k elements of el_size from which the mesher's
    ceil(length / el_size) gives k (k * el_size itself can round above k)."""
    L = k * el_size
    while np.ceil(L / el_size) > k:
        L = np.nextafter(L, 0.0)
    return L


def subdomain_block(n_dof_target, el_size=0.1e-3, rim=1e-3, n_src=64, rng=None, height=20e-3):
    """Steel rectangle of about n_dof_target DOF (DOF 1, 2), `height` tall
    and as wide as the DOF count asks, both whole multiples of el_size (so
    every block has the elements, the node-column height and the stable time
    step of the others), with an absorbing rim on every side, and n_src
    forcing/monitoring nodes spread along the inner top edge."""
    rng = np.random.default_rng(rng)
    rows = int(round(height / el_size))
    cols = max(int(round(n_dof_target / 2.0 / (rows + 1))) - 1, int(np.ceil(2.5 * rim / el_size)))
    H, W = _whole(rows, el_size), _whole(cols, el_size)
    matls = [steel()]
    mod = _typ(_m.fn_rectangular_structured_mesh(_rect(0, 0, W, H), el_size), matls)
    inner = _rect(rim, rim, W - rim, H - rim)
    mod = _m.fn_add_absorbing_layer(mod, inner, rim)
    opts = {"dof_to_use": [1, 2]}
    K, C, M, gl = _assemble(mod, matls, opts)
    xs = np.sort(rim + (W - 2 * rim) * rng.random(n_src))
    nd = _m.find_nodes_at_points(mod["nds"], np.stack([xs, np.full(n_src, H - rim)], axis=1))
    return dict(K=K, C=C, M=M, gl=gl, src=gl[nd - 1, 1] - 1, n=K.shape[0], mod=mod)


def union_mesh(mods):
    """One model holding several disjoint meshes, node and element numbers
    offset in order. With every node carrying the same DOFs its node-major
    DOF numbering is the meshes' numberings one after the other, so its
    assembled K, C, M are the block-diagonal matrices of the meshes."""
    offs = np.cumsum([0] + [m["nds"].shape[0] for m in mods])
    els = []
    for m, o in zip(mods, offs[:-1]):
        e = np.asarray(m["els"], dtype=np.int64).copy()
        e[e > 0] += o
        els.append(e)
    return {"nds": np.vstack([m["nds"] for m in mods]), "els": np.vstack(els),
            "el_mat_i": np.concatenate([np.asarray(m["el_mat_i"]) for m in mods]),
            "el_abs_i": np.concatenate([np.asarray(m.get("el_abs_i", np.zeros(len(m["el_mat_i"]))),
                                                   dtype=np.float64) for m in mods]),
            "el_typ_i": np.concatenate([np.asarray(m["el_typ_i"], dtype=object) for m in mods])}


def fmc_multimodel(n_el=64, pitch=0.3e-3, n_sub=8, sub_dof=(1e5, 1e6), n_t=3000, seed=4,
                   centre_freq=2e6, cycles=5, with_subdomains=True, plate=None, sub_height=20e-3):
    """Full-matrix capture on the 20 mm x 10 mm plate (64 elements at 0.3 mm
    pitch, one source column per transmitting element, every element
    receiving) plus n_sub subdomain models of seeded uneven size, each with 64
    source columns on its inner boundary, all marched as ONE block-diagonal
    system with 64 columns. Returns the combined K, C, M and 0-based forcing
    entries (dof, column) and monitored DOFs. ``plate`` overrides the plate's
    width and height (a dict, for reduced test models; the element size stays
    0.1 mm, that of the subdomain models); the subdomain models are
    ``sub_height`` tall and as wide as their seeded DOF counts ask."""
    import scipy.sparse as sp

    rng = np.random.default_rng(seed)
    main = array_plate(n_el=n_el, pitch=pitch, n_t=n_t, centre_freq=centre_freq, cycles=cycles,
                       **(plate or {}))
    blocks = [dict(K=main["K"], C=main["C"], M=main["M"], src=main["el_gi"] - 1, n=main["K"].shape[0],
                   mod=main["mod"])]
    if with_subdomains:
        sizes = rng.uniform(sub_dof[0], sub_dof[1], n_sub)
        for k in range(n_sub):
            blocks.append(subdomain_block(sizes[k], n_src=n_el, rng=rng, height=sub_height))
    offs = np.cumsum([0] + [b["n"] for b in blocks])
    K = sp.block_diag([b["K"] for b in blocks], format="csr")
    C = sp.block_diag([b["C"] for b in blocks], format="csr")
    M = sp.block_diag([b["M"] for b in blocks], format="csr")
    # same material and element size everywhere: the plate's stable step holds
    # for the subdomains too (their absorbing rims only lower the stiffness)
    dt = main["dt"]
    fdof, fsrc, mon = [], [], []
    for b, o in zip(blocks, offs[:-1]):
        fdof.append(b["src"] + o)
        fsrc.append(np.arange(n_el))
        mon.append(b["src"] + o)
    t = np.arange(n_t) * dt
    # the same system as one model of disjoint meshes (for the element operator)
    mod = union_mesh([b["mod"] for b in blocks])
    n_nodes = mod["nds"].shape[0]
    gl = np.zeros((n_nodes, 2), dtype=np.int64)
    gl[:, 0] = 2 * np.arange(n_nodes) + 1
    gl[:, 1] = 2 * np.arange(n_nodes) + 2
    assert 2 * n_nodes == K.shape[0]
    return dict(K=K, C=C, M=M, dt=dt, time=t, frcs=_m.hann_toneburst(t, centre_freq, cycles)[None, :],
                mod=mod, matls=[steel()], fe_options={"dof_to_use": [1, 2]}, gl_lookup=gl,
                frc_dof=np.concatenate(fdof), frc_src=np.concatenate(fsrc), hist_dof=np.concatenate(mon),
                n_src=n_el, model_sizes=[int(b["n"]) for b in blocks], offsets=offs,
                main_el_dof=main["el_gi"] - 1,
                name=(f"fmc {n_el} tx on plate_2d 20x10 mm h=0.1 mm (pitch {pitch*1e3:g} mm)"
                      + (f" + {n_sub} subdomain models" if with_subdomains else "")))

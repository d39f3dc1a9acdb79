"""Writes the committed fixtures of tests/golden/ (run from the repository
root: python tests/golden/make_golden.py).

Provenance: PROJECT ORACLE outputs, not reference outputs. The reference is
MATLAB-only and ships no golden vectors (its tests hold no assertions), so
these fixtures freeze what the oracle's literal restatement of
code/internal/fn_explicit_dynamic_solver_v6.m (oracle/v6.py, itself pinned by
closed forms and brute force in tests/test_oracle_pins.py) gives on small,
fully specified inputs:

  stability_mesh{0,1,2}  the tiny meshes of tests/stability.m:41-58 (one
                         steel element, one water element, the pair with an
                         ASI2D2 interface element), K and C dense; one
                         broadcast signal row whose first 9 samples are
                         zero, and velocity frames every 4 steps, so the
                         frames before the signal carry their times
                         (v6:163-166 starts a single-row march at ti = 1)
  coupled_small          aluminium under water with an interface and an
                         absorbing layer (CSR), two forcing rows with
                         leading zeros, 'KE' field frames every 50 steps
  tet_cube_small         a 4 mm C3D4 cube at 1 mm with side absorbing layers,
                         at half the CFL step (where both solver modes are
                         stable)

Each file holds the inputs (K, C, m = diag M, dt, n_t, forcing rows and
1-based DOFs, monitored 1-based DOFs) and the histories of both solver modes,
so a test needs neither the mesh generators nor the assembler to use it.
"""
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from conftest import PKG, assemble, coupled_block, steel, water  # noqa: E402
from oracle import assembly as oa  # noqa: E402
from oracle import v6 as ov6  # noqa: E402

MODES = ["vel at curent time step", "vel at last half time step"]
OUT = Path(__file__).resolve().parent
PROVENANCE = ("project oracle (oracle/v6.py, a restatement of "
              "code/internal/fn_explicit_dynamic_solver_v6.m); not reference output")


def _save(name, K, C, m, dt, n_t, fi, F, hi, field_every=0):
    K, C = sp.csr_matrix(K), sp.csr_matrix(C)
    t = np.arange(n_t) * dt
    rec = dict(k_ptr=K.indptr.astype(np.int64), k_col=K.indices.astype(np.int32), k_val=K.data,
               c_ptr=C.indptr.astype(np.int64), c_col=C.indices.astype(np.int32), c_val=C.data,
               m=np.asarray(m, float), dt=np.float64(dt), n_t=np.int64(n_t),
               frc_gi=np.asarray(fi, np.int64), frcs=np.atleast_2d(F).astype(float),
               hist_gi=np.asarray(hi, np.int64), field_every=np.int64(field_every),
               provenance=np.array(PROVENANCE))
    M = sp.diags(m)
    for k, mode in enumerate(MODES):
        every = field_every if field_every else np.inf
        h, f, ft, ti = ov6.explicit_dynamic_solver_v6(K, C, M, t, fi, F, hi, every, "KE", mode)
        rec[f"hist_mode{k}"] = h
        rec[f"ti_start_mode{k}"] = np.int64(ti)
        if field_every:
            rec[f"field_mode{k}"] = f
            rec[f"field_time_mode{k}"] = ft
    np.savez_compressed(OUT / f"{name}.npz", **rec)
    print(name, K.shape[0], "DOF,", K.nnz, "entries,", len(hi), "monitors x", n_t, "steps")


def stability_meshes():
    h = 1e-3
    s3 = np.sqrt(3)
    m1 = {"nds": np.array([[0.5, s3 / 2], [0, 0], [1, 0]]) * h, "els": np.array([[1, 2, 3]]),
          "el_mat_i": np.array([1]), "el_abs_i": np.zeros(1), "el_typ_i": ["CPE3"]}
    m2 = {"nds": np.array([[0, 0], [1, 0], [0.5, -s3 / 2]]) * h, "els": np.array([[1, 2, 3]]),
          "el_mat_i": np.array([2]), "el_abs_i": np.zeros(1), "el_typ_i": ["AC2D3"]}
    m3 = {"nds": np.array([[0.5, s3 / 2], [0, 0], [1, 0], [0.5, -s3 / 2]]) * h,
          "els": np.array([[1, 2, 3], [2, 3, 4]]), "el_mat_i": np.array([1, 2]),
          "el_abs_i": np.zeros(2), "el_typ_i": ["CPE3", "AC2D3"]}
    matls = [steel(), water()]
    for k, mod in enumerate((m1, m2, m3)):
        if k == 2:
            mod["el_typ_i"] = np.array(mod["el_typ_i"], dtype=object)
            mod = PKG.mesh.fn_add_fluid_solid_interface_els(mod, matls)
        K, C, M, _ = oa.build_global_matrices(mod["nds"], mod["els"], mod["el_mat_i"],
                                              mod["el_abs_i"], list(mod["el_typ_i"]), matls,
                                              {"dof_to_use": [1, 2, 4]})
        n = K.shape[0]
        m = M.diagonal()
        lam = np.max(np.abs(np.linalg.eigvals(np.diag(1 / m) @ K.toarray())))
        dt = 0.5 * 2 / np.sqrt(lam)
        n_t = 60
        s = np.arange(n_t - 9)
        F = np.concatenate([np.zeros(9), np.sin(s * 0.3 + 0.2) * np.exp(-0.05 * s)])
        _save(f"stability_mesh{k}", K, C, m, dt, n_t, [n], F, np.arange(1, n + 1), field_every=4)


def coupled_small():
    mod, matls = coupled_block(1.5e-3, 0.75e-3, 0.5e-3, 0.125e-3)
    K, C, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    dt = PKG.mesh.stable_time_step(K, M.diagonal())
    n_t = 200
    t = np.arange(n_t) * dt
    F = np.vstack([PKG.mesh.hann_toneburst(t - 20 * dt, 6e6, 3),
                   -0.5 * PKG.mesh.hann_toneburst(t - 35 * dt, 5e6, 2)])
    nd = PKG.mesh.fn_find_node_at_point(mod["nds"], [0.75e-3, 1.0e-3])
    both = np.nonzero((gl[:, 1] > 0) & (gl[:, 3] > 0))[0]
    fi = np.array([gl[nd - 1, 3], gl[both[len(both) // 2], 1]])
    hi = np.unique(np.concatenate([fi, np.arange(1, K.shape[0] + 1, 9)]))
    _save("coupled_small", K, C, M.diagonal(), dt, n_t, fi, F, hi, field_every=50)


def tet_cube_small():
    """At half the CFL step: with dt c/m above 2 in the layers at the full
    step, the 'vel at last half time step' recurrence (gamma = dt c/m - 1)
    diverges, and a fixture of two diverging runs pins nothing."""
    from bristolfe_v2_amd import configs
    c0 = configs.cube_3d(size=4e-3, el_size=1e-3, layer=1e-3, n_t=2, n_mon=4, mon_pitch=1e-3,
                         assembly_backend="numpy")
    c = configs.cube_3d(size=4e-3, el_size=1e-3, layer=1e-3, n_t=300, n_mon=4, mon_pitch=1e-3,
                        assembly_backend="numpy", dt=0.5 * c0["dt"])
    K, C, M = c["K"], c["C"], c["M"]
    hi = np.unique(np.concatenate([c["hist_gi"], np.arange(1, K.shape[0] + 1, 11)]))
    _save("tet_cube_small", K, C, M.diagonal(), c["dt"], 300, c["frc_gi"], c["frcs"], hi)


if __name__ == "__main__":
    stability_meshes()
    coupled_small()
    tet_cube_small()

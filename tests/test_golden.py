"""The committed fixtures of tests/golden/ (project-oracle outputs on small,
fully specified inputs; make_golden.py says how they were made): the
oracle still reproduces them (CPU), and the HIP march, called through the
C ABI as fn_explicit_dynamic_solver_v6, matches them to 1e-10 relative L2 on
every monitored history and 'KE' field frame, both solver modes (GPU)."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import PKG, rel_l2
from oracle import v6 as ov6

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["stability_mesh0", "stability_mesh1", "stability_mesh2", "coupled_small",
         "tet_cube_small"]
MODES = ["vel at curent time step", "vel at last half time step"]


def _load(name):
    z = np.load(GOLDEN / f"{name}.npz")  # plain arrays only (no pickles)
    n = z["m"].size
    K = sp.csr_matrix((z["k_val"], z["k_col"], z["k_ptr"]), shape=(n, n))
    C = sp.csr_matrix((z["c_val"], z["c_col"], z["c_ptr"]), shape=(n, n))
    t = np.arange(int(z["n_t"])) * float(z["dt"])
    every = int(z["field_every"]) or np.inf
    return z, K, C, sp.diags(z["m"]), t, every


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name, k):
    z, K, C, M, t, every = _load(name)
    assert "project oracle" in str(z["provenance"])
    h, f, ft, ti = ov6.explicit_dynamic_solver_v6(K, C, M, t, z["frc_gi"], z["frcs"],
                                                  z["hist_gi"], every, "KE", MODES[k])
    ref = z[f"hist_mode{k}"]
    assert np.linalg.norm(ref) > 0
    assert ti == int(z[f"ti_start_mode{k}"])
    assert rel_l2(h, ref) < 1e-13
    if np.isfinite(every):
        assert rel_l2(f, z[f"field_mode{k}"]) < 1e-13
        np.testing.assert_array_equal(ft, z[f"field_time_mode{k}"])


def test_single_row_fixtures_start_at_the_first_step():
    """The stability fixtures force with one broadcast row whose first 9
    samples are zero: v6:163-166 starts such a march at ti = 1, so the frames
    at ti = 1, 5, 9 hold zero velocity and carry their own times."""
    for name in CASES[:3]:
        z, K, C, M, t, every = _load(name)
        assert z["frcs"].shape[0] == 1 and not z["frcs"][0, :9].any() and z["frcs"][0, 9] != 0
        assert int(z["ti_start_mode0"]) == 1
        ft = z["field_time_mode0"]
        np.testing.assert_array_equal(ft, t[::every])
        assert not z["field_mode0"][:, :3].any() and z["field_mode0"][:, 3:].any()


def test_fixtures_cover_interface_damping_and_tets():
    z, K, C, M, t, every = _load("coupled_small")
    off = C - sp.diags(C.diagonal())
    assert off.nnz > 0 and C.diagonal().max() > 0  # interface couplings and absorbing damping
    assert z["frcs"].shape[0] == 2 and np.isfinite(every)
    z3, K3, *_ = _load("tet_cube_small")
    assert K3.shape[0] % 3 == 0 and K3.shape[0] > 300


@pytest.mark.gpu
@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("name", CASES)
def test_hip_march_matches_fixture(name, k):
    print("case: golden", name, MODES[k], flush=True)
    z, K, C, M, t, every = _load(name)
    h, f, _, ft = PKG.fn_explicit_dynamic_solver_v6(K, C, M, t, z["frc_gi"], z["frcs"], [], [],
                                                    z["hist_gi"], every, 1, "KE", MODES[k])
    assert rel_l2(h, z[f"hist_mode{k}"]) < 1e-10
    if np.isfinite(every):
        assert rel_l2(f, z[f"field_mode{k}"]) < 1e-10
        np.testing.assert_array_equal(ft, z[f"field_time_mode{k}"])

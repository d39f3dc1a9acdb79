"""Full-size checks on the GPU, at the benchmark workloads' own sizes, where
the oracle is too slow for whole marches: properties that hold exactly
whatever the size, and the oracle over the first steps.

* linearity: doubling the forcing doubles every history bit for bit (every
  operation of the march is linear, and scaling by 2 commutes with rounding
  in the normal range; the leading edge of a wave passes through subnormal
  values, about 1e-320, where it does not, so below 1e-290 the check is an
  absolute bound of 1e-300);
* time shift: delaying a single-row signal by k samples delays the histories
  by k steps bit for bit (v6 marches from ti = 1 and u stays exactly 0 until
  the signal starts, v6:163-166);
* two operators: the element-by-element and the CSR operator of the same
  plan (independent kernels and data) agree within the 1e-10 parity bar
  (about 1e-13 on the tets; about 3e-12 on the triangles, whose element
  scales carry J_rep / J_e and match K to 3e-13 per row);
* the oracle over the first steps: the oracle's C march (oracle/march.c,
  the v6 recurrence, OpenMP) runs the first 40 (cube) or 80 (coupled) steps
  of the full model, and both operators match it within 1e-10; field frames
  of the two operators agree within 1e-10 too, and repeated marches are
  bit-identical.

cube_3d: 24,361,803 DOF, 48M C3D4; coupled_2d: 7,677,395 DOF with interface
rows, absorbing layers and a void.
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle.march import OracleMarch

pytestmark = pytest.mark.gpu


def _workload(name, n_t):
    from bristolfe_v2_amd import configs, ebe
    c = getattr(configs, name)(n_t=n_t)
    K, C, M = c["K"], c["C"], c["M"]
    op = ebe.element_operator(c["mod"], c["matls"], c["fe_options"], c["gl_lookup"], K.shape[0])
    assert op is not None
    hi = np.unique(np.concatenate([c["hist_gi"] - 1, np.arange(0, K.shape[0], K.shape[0] // 61)]))
    return c, K, C, M, op, hi


@pytest.mark.parametrize("name,n_t", [("cube_3d", 300), ("coupled_2d", 600)])
def test_full_size_linearity_shift_and_operators(name, n_t):
    print("case: full size", name, n_t, flush=True)
    from bristolfe_v2_amd import _lib
    from bristolfe_v2_amd.solver import make_plan
    c, K, C, M, op, hi = _workload(name, n_t)
    F = np.atleast_2d(c["frcs"])
    fi = c["frc_gi"] - 1
    k = 37
    Fs = np.zeros_like(F)
    Fs[:, k:] = F[:, :-k]
    with make_plan(K, C, M, c["dt"], "vel at curent time step", element_op=op,
                   operator="ebe") as plan:
        assert plan.info()["operator_kind"] == 1
        h1, _, ti1 = plan.run(n_t, fi, F, hi)
        h2, _, _ = plan.run(n_t, fi, 2.0 * F, hi)
        hs, _, tis = plan.run(n_t, fi, Fs, hi)
        plan.set_tuning(_lib.TUNE_OPERATOR, _lib.OPERATOR_CSR)
        hc, _, _ = plan.run(n_t, fi, F, hi)
    assert ti1 == tis == 1                       # single-row forcing: v6 starts at ti = 1
    assert np.linalg.norm(h1) > 0 and np.isfinite(h1).all()
    normal = np.abs(h1) > 1e-290
    np.testing.assert_array_equal(h2[normal], 2.0 * h1[normal])
    assert np.abs(h2[~normal] - 2.0 * h1[~normal]).max(initial=0.0) <= 1e-300
    np.testing.assert_array_equal(hs[:, :k], 0.0)
    np.testing.assert_array_equal(hs[:, k:], h1[:, :-k])
    err = rel_l2(hc, h1)
    print(f"{name}: element operator vs CSR rel L2 {err:.3e}", flush=True)
    assert err < 1e-10


@pytest.mark.parametrize("name,n_t,oracle_steps", [("cube_3d", 300, 40), ("coupled_2d", 600, 80)])
def test_full_size_against_oracle_steps_and_frames(name, n_t, oracle_steps):
    print(f"case: full size vs oracle {name}, {n_t} steps", flush=True)
    from bristolfe_v2_amd import _lib, configs, ebe, solver
    c = getattr(configs, name)(n_t=n_t)
    K, C, M = c["K"], c["C"], c["M"]
    n = K.shape[0]
    assert n == {"cube_3d": 24_361_803, "coupled_2d": 7_677_395}[name]
    fi = np.asarray(c["frc_gi"]) - 1
    F = c["frcs"]
    # the configured monitors, the forcing DOFs and about a hundred rows
    # spread over the model
    hi = np.unique(np.concatenate([np.asarray(c["hist_gi"]) - 1, fi, np.arange(0, n, n // 97)]))
    op = ebe.element_operator(c["mod"], c["matls"], c["fe_options"], c["gl_lookup"], n)
    every = n_t - 1  # field frames at the first and the last step
    with solver.make_plan(K, C, M, c["dt"], element_op=op, operator="ebe") as plan:
        info = plan.info()
        assert info["operator_kind"] == _lib.OPERATOR_EBE
        assert info["ebe_check"] < 1e-11
        h_e, f_e, ti = plan.run(n_t, fi, F, hi, every, 1)
        h_e2, _, _ = plan.run(n_t, fi, F, hi)
        plan.set_tuning(_lib.TUNE_OPERATOR, _lib.OPERATOR_CSR)
        assert plan.info()["operator_kind"] == _lib.OPERATOR_CSR
        h_c, f_c, ti_c = plan.run(n_t, fi, F, hi, every, 1)
    print(f"  ebe vs csr: histories {rel_l2(h_e, h_c):.2e}, frames {rel_l2(f_e, f_c):.2e}",
          flush=True)
    assert ti == ti_c == 1
    assert np.all(np.isfinite(h_e)) and np.linalg.norm(h_e) > 0
    assert np.linalg.norm(f_e[:, -1]) > 0
    np.testing.assert_array_equal(h_e, h_e2)
    assert rel_l2(h_e, h_c) < 1e-10
    assert rel_l2(f_e, f_c) < 1e-10
    ref, rti = OracleMarch(K, C, M, c["dt"]).run(n_t, fi, F, hi, t_end=oracle_steps)
    assert rti == 1
    s = oracle_steps
    print(f"  vs oracle over {s} steps: ebe {rel_l2(h_e[:, :s], ref[:, :s]):.2e}, "
          f"csr {rel_l2(h_c[:, :s], ref[:, :s]):.2e}", flush=True)
    assert np.linalg.norm(ref[:, :s]) > 0
    assert rel_l2(h_e[:, :s], ref[:, :s]) < 1e-10
    assert rel_l2(h_c[:, :s], ref[:, :s]) < 1e-10

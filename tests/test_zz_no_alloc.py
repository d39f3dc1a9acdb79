"""bfe_disps_to_forces_noalloc (the allocation-free variant of
bfe_disps_to_forces, all device arrays in the library's static 64 MiB scratch
array) against bfe_disps_to_forces, through the C ABI: two calls in a row on
the same GPU, the second smaller than the first (the scratch keeps the first
call's contents), both modes and both sides; and the size limit, which is
checked before any device call."""
import ctypes as C

import numpy as np
import pytest

from conftest import PKG, assemble, coupled_block, solid_plate

pytestmark = pytest.mark.gpu

MODES = ["vel at curent time step", "vel at last half time step"]


def _inputs(model, n_t, seed):
    if model == "coupled":
        mod, matls = coupled_block(1.5e-3, 0.75e-3, 0.5e-3, 0.125e-3)
        K, Cm, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2, 4], "damping_power_law": 2})
    else:
        mod, matls = solid_plate(0.8e-3, 0.4e-3, 0.1e-3)
        K, Cm, M, gl = assemble(mod, matls, {"dof_to_use": [1, 2]})
    n = K.shape[0]
    rng = np.random.default_rng(seed)
    disp = rng.standard_normal((n, n_t))
    lyrs = rng.integers(0, 5, n)
    return K, Cm, np.asarray(M.sum(axis=0)).ravel(), disp, lyrs


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("side", [0, 1])
def test_noalloc_matches_allocating_entry_point(mode, side):
    from bristolfe_v2_amd import _lib
    print("case: disps_to_forces_noalloc", MODES[mode], "in" if side == 0 else "out", flush=True)
    for model, n_t, seed in (("coupled", 120, 1), ("plate", 37, 2)):  # second call smaller
        K, Cm, m, disp, lyrs = _inputs(model, n_t, seed)
        i0, f0 = _lib.disps_to_forces(K, Cm, m, 1e-8, disp, lyrs, side, mode)
        i1, f1 = _lib.disps_to_forces(K, Cm, m, 1e-8, disp, lyrs, side, mode, no_alloc=True)
        np.testing.assert_array_equal(i1, i0)
        assert f0.shape == f1.shape and np.abs(f0).max() > 0
        np.testing.assert_array_equal(f1, f0)


def test_noalloc_rejects_inputs_beyond_its_scratch():
    from bristolfe_v2_amd import _lib
    K, Cm, m, _, lyrs = _inputs("plate", 1, 3)
    n = K.shape[0]
    n_t = (64 << 20) // (8 * n) + 1            # the displacements alone exceed 64 MiB
    disp = np.zeros((n, n_t), order="F")
    with pytest.raises(_lib.BfeError) as e:
        _lib.disps_to_forces(K, Cm, m, 1e-8, disp, lyrs, 0, 0, no_alloc=True)
    assert e.value.status == _lib.BFE_ERR_ARG and "bfe_disps_to_forces" in str(e.value)
    # the allocating entry point takes the same call
    idx, f = _lib.disps_to_forces(K, Cm, m, 1e-8, disp, lyrs, 0, 0)
    assert f.shape == (idx.size, n_t)

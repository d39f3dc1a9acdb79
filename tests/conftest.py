import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

PKG = __graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def pkg():
    return PKG


@pytest.fixture(scope="session")
def mesh():
    return PKG.mesh


def steel():
    return {"rho": 8000.0, "D": PKG.mesh.fn_isotropic_stiffness_matrix(210e9, 0.3), "el_typ": "CPE3",
            "name": "Steel"}


def aluminium():
    return {"rho": 2700.0, "D": PKG.mesh.fn_isotropic_stiffness_matrix(70e9, 0.33), "el_typ": "CPE3",
            "name": "Aluminium"}


def water():
    return {"rho": 1000.0, "D": 1480.0 ** 2 * 1000.0, "el_typ": "AC2D3", "name": "Water"}


def rect(w, h, x0=0.0, y0=0.0):
    return np.array([[x0, y0], [x0 + w, y0], [x0 + w, y0 + h], [x0, y0 + h]])


def solid_plate(w=4e-3, h=2e-3, el=0.2e-3, absorbing=None):
    m = PKG.mesh
    mod = m.fn_rectangular_structured_mesh(rect(w, h), el)
    if absorbing is not None:
        mod = m.fn_add_absorbing_layer(mod, absorbing[0], absorbing[1])
    mod["el_typ_i"] = np.array(["CPE3"] * len(mod["el_mat_i"]), dtype=object)
    return mod, [steel()]


def coupled_block(w=3e-3, hs=1.5e-3, hf=1.0e-3, el=0.25e-3, absorbing=True, void=None):
    """Aluminium block under a water layer, interface elements, optional
    absorbing layer on left/right/bottom and an optional void polygon."""
    m = PKG.mesh
    matls = [aluminium(), water()]
    mod = m.fn_rectangular_structured_mesh(rect(w, hs + hf), el)
    mod = m.fn_set_els_inside_bdry_to_mat(mod, rect(w + 2, hf, -1, hs), 2)
    mod["el_typ_i"] = np.asarray([matls[i - 1]["el_typ"] for i in mod["el_mat_i"]], dtype=object)
    if void is not None:
        mod = m.remove_elements_in_region(mod, void)
    if absorbing:
        t = 0.2 * w
        inner = np.array([[t, t], [w - t, t], [w - t, 10 * (hs + hf)], [t, 10 * (hs + hf)]])
        mod = m.fn_add_absorbing_layer(mod, inner, t)
    mod = m.fn_add_fluid_solid_interface_els(mod, matls)
    return mod, matls


def assemble(mod, matls, opts=None):
    return PKG.fn_build_global_matrices_v5(mod["nds"], mod["els"], mod["el_mat_i"], mod["el_abs_i"],
                                           mod["el_typ_i"], matls, opts or {})


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0)

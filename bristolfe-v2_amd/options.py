"""Option defaults, mirroring the layered struct defaults of the reference.

fn_set_default_fields (code/internal/fn_set_default_fields.m:19-25) adds a
default only where the field is absent; existing values are kept.
"""
from __future__ import annotations

# code/fn_FE_entry_point.m:23-35
ENTRY_POINT_DEFAULTS = {
    "solver": "BristolFE",
    "damping_power_law": 3,
    "max_damping": 3.1415e7,
    "max_stiffness_reduction": 0.01,
    "solver_precision": "double",
    "field_output_every_n_frames": float("inf"),
    "dof_to_use": [],
}

# code/fn_BristolFE_v2.m:22-30
BRISTOLFE_DEFAULTS = {
    "use_gpu_if_available": 1,
    "global_matrix_builder_version": "v5",
    "dynamic_solver_version": "v6",
    "solver_mode": "vel at curent time step",  # spelling as in the reference
    "field_output_type": "KE",
    # not in the reference: K u by the element-by-element operator when the
    # mesh has the template form ("auto"), always CSR ("csr"), or required ("ebe")
    "stiffness_operator": "auto",
    # not in the reference: the 'KE' field output reduced per element on the
    # device next to the march (True) or on the host from the frames (False)
    "device_field_output": True,
}

SOLVER_MODES = {
    "vel at curent time step": 0,     # reference spelling (fn_BristolFE_v2.m:28)
    "vel at current time step": 0,    # correct spelling, accepted as an alias
    "vel at last half time step": 1,
}


def set_default_fields(old: dict | None, defaults: dict) -> dict:
    new = dict(old or {})
    for k, v in defaults.items():
        if k not in new:
            new[k] = list(v) if isinstance(v, list) else v
    return new

"""BristolFE-v2 explicit time-march on MI355X (gfx950).

The directory name is not a Python identifier; entry points register it as
the package ``bristolfe_v2_amd`` with :func:`load_package` semantics (see
``__graft_entry__.py``). Public API:

  fn_FE_entry_point, fn_BristolFE_v2           (driver)
  fn_explicit_dynamic_solver_v6                (march, v6 signature)
  fn_build_global_matrices_v5                  (assembly)
  fn_get_field_output                          (field post-processing)
  mesh / signals helpers in ``mesh``
"""
from .assembly import fn_build_global_matrices_v5  # noqa: F401
from .driver import fn_BristolFE_v2, fn_FE_entry_point, fn_nds_and_dfs_to_gi  # noqa: F401
from .field_output import fn_get_field_output  # noqa: F401
from .solver import fn_explicit_dynamic_solver_v6, make_plan  # noqa: F401
from . import mesh  # noqa: F401
from . import subdoms  # noqa: F401

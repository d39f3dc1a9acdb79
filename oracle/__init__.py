"""ORACLE — TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference BristolFE-v2 algorithm for the explicit
time-march path. It is the checker for the HIP product path and the CPU
baseline of bench.py. Only tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg may import, call or link anything here; the product package
never does.

Modules
  v6.py        literal restatement of fn_explicit_dynamic_solver_v6.m: forms
               A, B, D exactly as v6:130-139 (``X \\ Y`` by block Gaussian
               elimination) and runs u = A f + B u1 + D u2 per step.
  elements.py  per-element closed forms of fn_el_CPE3/AC2D3/ASI2D2.m and the
               new C3D4, written with explicit loops.
  assembly.py  per-element loop restatement of fn_build_global_matrices_v5.m.
  post.py      fn_get_field_output.m ('KE', means) and
               tests/fn_amplification_matrix.m (stability).
  march.c      C/OpenMP fused march (same algebra as v6 after eliminating the
               diagonal X), used for the CPU timing and large-size checks.

Pins (tests/test_oracle_pins.py): SDOF closed form, dense brute force on the
meshes of tests/stability.m, mode equivalence at C = 0, reciprocity,
discrete energy conservation, ti_start exactness, time-shift and linearity,
element rigid-body modes, patch test, symmetry, quadrature cross-check.
Parity against MATLAB itself is unpinned: MATLAB is not available, and the
reference ships no golden vectors (its tests contain no assertions).
"""

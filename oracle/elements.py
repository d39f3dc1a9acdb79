"""Per-element closed forms, one element at a time.

ORACLE — test infrastructure only (see oracle/__init__.py).

  cpe3(xy, D, rho)   code/internal/elements/fn_el_CPE3.m:46-158
      J = x1 y2 - y1 x2 - x1 y3 + y1 x3 + x2 y3 - y2 x3           (:56)
      strain rows (e11, e22, e33, g23, g13, g12), d/dz = 0
      (el_builder/el_sym_CPE3.m:58-64); Ke = (J/2) B^T D B; M = J rho/6 I
  ac2d3(xy, K, rho)  code/internal/elements/fn_el_AC2D3.m:46-80
      Ke = -(1/rho) (grad N)^T (grad N) (J/2); M = -J/(6 K) I
  asi2d2(xy)         code/internal/elements/fn_el_ASI2D2.m:46-82
      C(u1_a, p_b) = (y2 - y1)/4, C(u2_a, p_b) = (x1 - x2)/4, symmetric
  c3d4(xyz, D, rho)  no counterpart: Ke = (J/6) B^T D B, M = J rho/24 I
Local DOF order is node-major with DOFs (1, 2, 3) for CPE3 and C3D4, (4) for
AC2D3 and (1, 2, 3, 4) for ASI2D2.
"""
from __future__ import annotations

import numpy as np


def _tri_grads(xy):
    (x1, y1), (x2, y2), (x3, y3) = xy
    J = x1 * y2 - y1 * x2 - x1 * y3 + y1 * x3 + x2 * y3 - y2 * x3
    # gradient of the linear shape functions: dN_i/dx = b_i/J, dN_i/dy = c_i/J
    b = [y2 - y3, y3 - y1, y1 - y2]
    c = [x3 - x2, x1 - x3, x2 - x1]
    return J, [bi / J for bi in b], [ci / J for ci in c]


def cpe3(xy, D, rho):
    J, dx, dy = _tri_grads(xy)
    B = np.zeros((6, 9))
    for i in range(3):
        ux, uy, uz = 3 * i, 3 * i + 1, 3 * i + 2
        B[0, ux] = dx[i]
        B[1, uy] = dy[i]
        B[3, uz] = dy[i]
        B[4, uz] = dx[i]
        B[5, ux] = dy[i]
        B[5, uy] = dx[i]
    K = np.zeros((9, 9))
    for a in range(9):
        for b in range(9):
            s = 0.0
            for r in range(6):
                for t in range(6):
                    s += B[r, a] * D[r][t] * B[t, b]
            K[a, b] = s * J / 2.0
    M = np.eye(9) * (J * rho / 6.0)
    return K, np.zeros((9, 9)), M, [1, 1, 1, 2, 2, 2, 3, 3, 3], [1, 2, 3, 1, 2, 3, 1, 2, 3]


def ac2d3(xy, bulk, rho):
    J, dx, dy = _tri_grads(xy)
    K = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            K[i, j] = -(dx[i] * dx[j] + dy[i] * dy[j]) * J / 2.0 / rho
    M = np.eye(3) * (-J / (6.0 * bulk))
    return K, np.zeros((3, 3)), M, [1, 2, 3], [4, 4, 4]


def asi2d2(xy):
    (x1, y1), (x2, y2) = xy
    C = np.zeros((8, 8))
    for a in (0, 4):
        for pb in (3, 7):
            for d, v in ((0, (y2 - y1) / 4.0), (1, (x1 - x2) / 4.0)):
                C[a + d, pb] = v
                C[pb, a + d] = v
    Z = np.zeros((8, 8))
    return Z, C, Z.copy(), [1, 1, 1, 1, 2, 2, 2, 2], [1, 2, 3, 4, 1, 2, 3, 4]


def c3d4(xyz, D, rho):
    X = np.asarray(xyz, dtype=np.float64)
    E = np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]])
    J = np.linalg.det(E)
    Jm = E.T  # Jm[j, k] = dx_j / dxi_k
    inv = np.linalg.inv(Jm)  # inv[k, j] = dxi_k / dx_j
    gref = [[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    B = np.zeros((6, 12))
    for i in range(4):
        g = [sum(gref[i][k] * inv[k, j] for k in range(3)) for j in range(3)]
        ux, uy, uz = 3 * i, 3 * i + 1, 3 * i + 2
        B[0, ux] = g[0]
        B[1, uy] = g[1]
        B[2, uz] = g[2]
        B[3, uy] = g[2]
        B[3, uz] = g[1]
        B[4, ux] = g[2]
        B[4, uz] = g[0]
        B[5, ux] = g[1]
        B[5, uy] = g[0]
    K = B.T @ np.asarray(D) @ B * (J / 6.0)
    M = np.eye(12) * (J * rho / 24.0)
    return K, np.zeros((12, 12)), M, [1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4], [1, 2, 3] * 4


def element(typ, coords, D, rho):
    if typ == "CPE3":
        return cpe3(coords, D, rho)
    if typ == "AC2D3":
        return ac2d3(coords, float(np.asarray(D).ravel()[0]), rho)
    if typ == "ASI2D2":
        return asi2d2(coords)
    if typ == "C3D4":
        return c3d4(coords, D, rho)
    raise ValueError(typ)


NODES_PER_EL = {"CPE3": 3, "AC2D3": 3, "ASI2D2": 2, "C3D4": 4}

"""ORACLE -- test infrastructure only (imported by tests/ as the checker; never by the
product).

NumPy restatement of the Gauss-point gradient fields of the scalar Q4 mesh behind the
virtual-heat time fields (sttopt/virtual_heat.py:51-72 and :290-347):

- shape_gradients: the bilinear shape functions N_BL = (1-xi)(1-eta), N_BR = xi(1-eta),
  N_TR = xi eta, N_TL = (1-xi) eta (local node order of sttopt/fem.py:89-95, x right,
  y up) differentiated at the 2x2 Gauss points, eta-major (virtual_heat.py:51-72);
- gauss_gradients: G[e, g, d] = sum_i B[g, d, i] u[edof[e, i]] (virtual_heat.py:290-292);
- direction_field: G / sqrt(|G|^2 + tiny) (virtual_heat.py:294-307);
- divergence_load: F_i = 1/4 sum_e w_e sum_{g,d} X[e, g, d] B[g, d, loc_i(e)], scattered
  element by element (virtual_heat.py:309-315);
- poisson_distance: the assembled direct solve of div(w grad phi) = div(w X), phi = 0
  on the plate (virtual_heat.py:317-347).

Pinned by tests/test_oracle.py: the shape gradients against central differences of
the shape functions and partition of unity; the gradients of the coordinate fields x
and y and of the bilinear field x y (exact at the Gauss points, which fixes the axis
directions and the Gauss-point order); the load of a constant field against its
boundary-flux closed form; and the load against the gradients through the identity
F . phi = 1/4 sum_e w_e X_e : G(phi)_e.
"""

from __future__ import annotations

import numpy as np

from . import fem_direct

TINY = np.finfo(np.float64).tiny


def gauss_points() -> list[tuple[float, float]]:
    """(xi, eta) of the 2x2 rule on the unit square, eta-major."""
    a = 0.5 / np.sqrt(3.0)
    g = (0.5 - a, 0.5 + a)
    return [(g[i % 2], g[i // 2]) for i in range(4)]


def shape_functions(xi: float, eta: float) -> np.ndarray:
    return np.array([(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta])


def shape_gradients() -> np.ndarray:
    """B[g, d, i] = dN_i/dx_d at Gauss point g (element length 1)."""
    # N_i = X_i(xi) Y_i(eta) with X = (1 - xi) or xi and Y = (1 - eta) or eta
    sx = np.array([-1.0, 1.0, 1.0, -1.0])  # d/dxi of the xi factor
    sy = np.array([-1.0, -1.0, 1.0, 1.0])  # d/deta of the eta factor
    B = np.zeros((4, 2, 4))
    for g, (xi, eta) in enumerate(gauss_points()):
        fx = np.where(sx > 0, xi, 1 - xi)
        fy = np.where(sy > 0, eta, 1 - eta)
        B[g, 0] = sx * fy
        B[g, 1] = fx * sy
    return B


def gauss_gradients(u: np.ndarray, nelx: int, nely: int) -> np.ndarray:
    """(nel, 4, 2) gradients of the nodal field u at the Gauss points."""
    edof = fem_direct.element_dofs(nelx, nely, 1)
    B = shape_gradients()
    return np.einsum("gdi,ei->egd", B, u[edof])


def direction_field(s: np.ndarray, nelx: int, nely: int) -> np.ndarray:
    G = gauss_gradients(s, nelx, nely)
    return G / np.sqrt((G**2).sum(-1) + TINY)[..., None]


def divergence_load(w: np.ndarray, X: np.ndarray, nelx: int, nely: int) -> np.ndarray:
    """F_i = 1/4 sum_e w_e X_e . grad N_i at the Gauss points, element by element."""
    edof = fem_direct.element_dofs(nelx, nely, 1)
    B = shape_gradients()
    F = np.zeros((nelx + 1) * (nely + 1))
    for e in range(nelx * nely):
        per_node = w[e] * np.einsum("gd,gdi->i", X[e], B) / 4
        np.add.at(F, edof[e], per_node)
    return F


def poisson_distance(s: np.ndarray, w: np.ndarray, plate: np.ndarray, nelx: int,
                     nely: int) -> np.ndarray:
    """phi with div(w grad phi) = div(w X(s)) by the assembled direct solve, phi = 0 on
    the plate nodes."""
    KE = fem_direct.diffusion_KE()
    edof = fem_direct.element_dofs(nelx, nely, 1)
    ndof = (nelx + 1) * (nely + 1)
    K = fem_direct.assemble(KE, w, edof, ndof)
    F = divergence_load(w, direction_field(s, nelx, nely), nelx, nely)
    return fem_direct.lifted_solve(K, F, np.zeros(ndof), plate)

"""Mesh and element setup for the structured Q4 grid (the inputs of the solve).

Same conventions as the reference (sttopt/fem.py:60-95, sttopt/conventions.md): an
`nelx` x `nely` grid of unit-square bilinear quads, fields shaped `(nely, nelx)` in C
order with row 0 at the top, node `(r, c)` numbered `r * (nelx + 1) + c`, `d` dofs per
node interleaved, element-local corners bottom-left, bottom-right, top-right,
top-left.  The element matrices are computed here by 2 x 2 Gauss quadrature of the
bilinear shape functions rather than typed in; a unit test pins them against the
closed-form table the reference uses (sttopt/fem.py:17-30, 98-111).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

# Corner (xi, eta) of each element-local node, in the order BL, BR, TR, TL
# (x to the right, y up).
_CORNERS = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])
_GAUSS = np.array([-1.0, 1.0]) / np.sqrt(3.0)


def _shape_gradients(xi: float, eta: float) -> np.ndarray:
    """d N_i / d(x, y) on the unit square at natural point (xi, eta): shape (4, 2)."""
    g = np.empty((4, 2))
    for i, (xi_i, eta_i) in enumerate(_CORNERS):
        dxi = xi_i * (1 + eta * eta_i) / 4
        deta = eta_i * (1 + xi * xi_i) / 4
        g[i] = (2 * dxi, 2 * deta)  # x = (1 + xi) / 2, so d/dx = 2 d/dxi
    return g


def plane_stress_KE(nu: float) -> np.ndarray:
    """8x8 plane-stress stiffness of a unit-square Q4 element with E = 1.

    Dofs (x, y) per node, nodes BL, BR, TR, TL.  Full 2 x 2 Gauss integration of
    B^T D B with D = [[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]] / (1 - nu^2).
    """
    D = np.array([[1.0, nu, 0.0], [nu, 1.0, 0.0], [0.0, 0.0, (1.0 - nu) / 2]])
    D /= 1.0 - nu**2
    KE = np.zeros((8, 8))
    for xi in _GAUSS:
        for eta in _GAUSS:
            g = _shape_gradients(xi, eta)
            B = np.zeros((3, 8))
            B[0, 0::2] = g[:, 0]
            B[1, 1::2] = g[:, 1]
            B[2, 0::2] = g[:, 1]
            B[2, 1::2] = g[:, 0]
            KE += B.T @ D @ B * 0.25  # |J| = 1/4, unit Gauss weights
    return (KE + KE.T) / 2


def diffusion_KE() -> np.ndarray:
    """4x4 element matrix of -div(grad u) on a unit-square Q4 element (nodes BL, BR, TR, TL)."""
    KE = np.zeros((4, 4))
    for xi in _GAUSS:
        for eta in _GAUSS:
            g = _shape_gradients(xi, eta)
            KE += g @ g.T * 0.25
    return (KE + KE.T) / 2


def node_grid(nelx: int, nely: int) -> np.ndarray:
    """Global node number at each corner position `(row, col)`, shape `(nely+1, nelx+1)`."""
    return np.arange((nely + 1) * (nelx + 1)).reshape(nely + 1, nelx + 1)


def element_dof_map(nelx: int, nely: int, dofs_per_node: int = 2) -> np.ndarray:
    """Global dofs of each element's corners (order BL, BR, TR, TL), shape (nel, 4d)."""
    n = node_grid(nelx, nely)
    tl = n[:-1, :-1].ravel()
    corners = np.stack([tl + nelx + 1, tl + nelx + 2, tl + 1, tl], axis=1)
    d = dofs_per_node
    return (d * corners[:, :, None] + np.arange(d)).reshape(-1, 4 * d)


def edge_load_shares(n_nodes: int, length: float) -> np.ndarray:
    """Nodal shares of a unit force spread uniformly over `[0, length]` elements of an
    edge, measured from the edge's first node (exact hat-function integrals; a
    zero length is a point load on the first node)."""
    if not 0 <= length <= n_nodes - 1:
        raise ValueError(f"a traction over {length} elements does not fit on {n_nodes - 1}")
    if length == 0:
        out = np.zeros(n_nodes)
        out[0] = 1.0
        return out

    def cumulative(s):
        s = np.clip(s, -1.0, 1.0)
        return np.where(s < 0, (1 + s) ** 2 / 2, 1 - (1 - s) ** 2 / 2)

    i = np.arange(n_nodes)
    return (cumulative(length - i) - cumulative(-i)) / length


def gravity_load_matrix(nelx: int, nely: int) -> sp.csr_matrix:
    """C with C[node, e] = fe / 4 at the element's four corners, fe = 1 / (nelx nely):
    `C @ x.ravel()` is the nodal self-weight of density field x at unit total weight
    (sttopt/gravity.py:15-36)."""
    nel = nelx * nely
    edof = element_dof_map(nelx, nely, 1)
    rows = edof.ravel()
    cols = np.repeat(np.arange(nel), 4)
    vals = np.full(4 * nel, 1.0 / nel / 4)
    return sp.coo_matrix((vals, (rows, cols)), shape=((nelx + 1) * (nely + 1), nel)).tocsr()


def load_case(case: str, nelx: int, nely: int, load_length: float = 0.0,
              support_length: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
    """`(F, freedofs)` of the structural cases the solve is exercised on, lengths in
    elements (sttopt/load_cases.py:74-130 for "cantilever" and "mbb").

    - "cantilever": left edge clamped, unit -y load at the bottom-right corner
      (spread upward over `load_length`).
    - "mbb": half MBB beam -- x fixed on the left edge, y fixed on a support patch of
      `support_length` from the bottom-right corner, unit -y load at the top-left
      corner (spread right over `load_length`).
    - "bottom_clamped": bottom edge clamped, no external load (self-weight cases).
    - "bottom_clamped_tr": bottom edge clamped, unit -y load at the top-right corner.
    """
    nodes = node_grid(nelx, nely)
    ndof = 2 * nodes.size
    F = np.zeros(ndof)
    if case == "cantilever":
        right_up = nodes[::-1, -1]
        F[2 * right_up + 1] = -edge_load_shares(nely + 1, load_length)
        fixed = np.stack([2 * nodes[:, 0], 2 * nodes[:, 0] + 1], axis=1).ravel()
    elif case == "mbb":
        F[2 * nodes[0, :] + 1] = -edge_load_shares(nelx + 1, load_length)
        bottom_from_right = nodes[-1, ::-1]
        pos = np.arange(bottom_from_right.size)
        roller = bottom_from_right[pos <= support_length + 1e-9]
        fixed = np.concatenate([2 * nodes[:, 0], 2 * roller + 1])
    elif case in ("bottom_clamped", "bottom_clamped_tr"):
        bottom = nodes[-1, :]
        fixed = np.stack([2 * bottom, 2 * bottom + 1], axis=1).ravel()
        if case == "bottom_clamped_tr":
            F[2 * nodes[0, -1] + 1] = -1.0
    else:
        raise ValueError(f"unknown load case {case!r}")
    return F, np.setdiff1d(np.arange(ndof), fixed)

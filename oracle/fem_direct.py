"""ORACLE -- test infrastructure only (imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg as the checker; never by the product).

Assembled-matrix fp64 oracle of the solve-and-sensitivity path, restated with
NumPy/SciPy from the reference's independent test oracles:

- element matrices: the closed-form tables of sttopt/fem.py:17-30 and :98-111
  (an independent derivation from the product's Gauss-quadrature ones);
- dof map: sttopt/fem.py:60-95 written as explicit loops;
- assembly and direct solve: tests/reference/fem.py:16-70 (COO assembly,
  symmetrized; `spsolve` on the free-dof block);
- self-weight matrix: sttopt/gravity.py:15-36;
- compliance and hand-derived sensitivities: tests/reference/compliance.py:24-110
  (whole structure: dc/dx = -p dE x^(p-1) c_e; stage: the SIMP term times the mask
  plus the doubled gravity-load adjoint term -2 (C^T U_y)_e).

Pinned by tests/test_oracle.py against closed forms (patch tests, a bar in
tension, a clamped-column self-weight reaction), brute-force dense assembly, the
MATLAB-style MBB numbers, and central finite differences.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def plane_stress_KE(nu: float) -> np.ndarray:
    """The reference's closed-form 8x8 table (sttopt/fem.py:17-30)."""
    A11 = np.array([[12, 3, -6, -3], [3, 12, 3, 0], [-6, 3, 12, -3], [-3, 0, -3, 12]])
    A12 = np.array([[-6, -3, 0, 3], [-3, -6, -3, -6], [0, -3, -6, 3], [3, -6, 3, -6]])
    B11 = np.array([[-4, 3, -2, 9], [3, -4, -9, 4], [-2, -9, -4, -3], [9, 4, -3, -4]])
    B12 = np.array([[2, -3, 4, -9], [-3, 2, 9, -2], [4, 9, 2, 3], [-9, -2, 3, 2]])
    A = np.block([[A11, A12], [A12.T, A11]])
    B = np.block([[B11, B12], [B12.T, B11]])
    return (A + nu * B) / (1 - nu**2) / 24


def diffusion_KE() -> np.ndarray:
    """The reference's closed-form 4x4 table (sttopt/fem.py:98-111)."""
    return np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]]) / 6


def element_dofs(nelx: int, nely: int, d: int = 2) -> np.ndarray:
    """Element dof map by explicit loops: element (r, c), corners BL, BR, TR, TL."""
    out = np.empty((nelx * nely, 4 * d), dtype=np.int64)
    for r in range(nely):
        for c in range(nelx):
            tl = r * (nelx + 1) + c
            corners = (tl + nelx + 1, tl + nelx + 2, tl + 1, tl)
            out[r * nelx + c] = [d * n + k for n in corners for k in range(d)]
    return out


def assemble(KE: np.ndarray, density: np.ndarray, edof: np.ndarray, ndof: int,
             reaction: np.ndarray | None = None) -> sp.csr_matrix:
    """Global K = sum_e density_e KE (+ reaction_e I/4 lumped), symmetrized."""
    k = KE.shape[0]
    rows = np.repeat(edof, k, axis=1).ravel()
    cols = np.tile(edof, (1, k)).ravel()
    vals = (density[:, None, None] * KE[None]).reshape(len(density), -1)
    K = sp.coo_matrix((vals.ravel(), (rows, cols)), shape=(ndof, ndof)).tocsr()
    if reaction is not None:
        diag = np.zeros(ndof)
        np.add.at(diag, edof.ravel(), np.repeat(reaction / 4, k))
        K = K + sp.diags(diag)
    return ((K + K.T) / 2).tocsr()


def solve(K: sp.spmatrix, F: np.ndarray, freedofs: np.ndarray) -> np.ndarray:
    """Direct solve of the free-dof block; fixed dofs 0."""
    U = np.zeros(K.shape[0])
    Kf = K[np.ix_(freedofs, freedofs)].tocsc()
    U[freedofs] = spla.spsolve(Kf, F[freedofs])
    return U


def lifted_solve(K: sp.spmatrix, F: np.ndarray, g: np.ndarray, fixed: np.ndarray) -> np.ndarray:
    """u with u = g on `fixed` (bool) and K u = F on the free dofs."""
    free = np.flatnonzero(~fixed)
    lift = np.where(fixed, g, 0.0)
    rhs = F - K @ lift
    u = lift.copy()
    u[free] = spla.spsolve(K[np.ix_(free, free)].tocsc(), rhs[free])
    return u


def strain_energy(U: np.ndarray, edof: np.ndarray, KE: np.ndarray) -> np.ndarray:
    Ue = U[edof]
    return np.einsum("ei,ij,ej->e", Ue, KE, Ue)


def simp(x, emin, emax, p):
    return emin + x**p * (emax - emin)


def gravity_matrix(nelx: int, nely: int) -> sp.csr_matrix:
    """C[node, e] = fe / 4 on the element's corners, fe = 1 / (nelx nely)."""
    nel = nelx * nely
    corners = element_dofs(nelx, nely, 1)
    rows = corners.ravel()
    cols = np.repeat(np.arange(nel), 4)
    return sp.coo_matrix((np.full(4 * nel, 0.25 / nel), (rows, cols)),
                         shape=((nelx + 1) * (nely + 1), nel)).tocsr()


def time_mask(t: np.ndarray, ti: float, beta: float) -> np.ndarray:
    num = math.tanh(beta * ti) + np.tanh(beta * (t - ti))
    den = math.tanh(beta * ti) + math.tanh(beta * (1 - ti))
    return np.maximum(1 - num / den, 0.0)


def time_mask_derivative(t: np.ndarray, ti: float, beta: float) -> np.ndarray:
    num = beta * (np.tanh(beta * (t - ti)) ** 2 - 1)
    den = math.tanh(beta * (ti - 1)) - math.tanh(beta * ti)
    out = -num / den
    # the clamp of time_mask zeroes the derivative where it acts
    return np.where(1 - (math.tanh(beta * ti) + np.tanh(beta * (t - ti)))
                    / (math.tanh(beta * ti) + math.tanh(beta * (1 - ti))) >= 0, out, 0.0)


def whole_compliance(x, KE, edof, F, freedofs, emin=1e-9, emax=1.0, p=3.0):
    """`(c, dc/dx, U)` under a fixed load (tests/reference/compliance.py:33-48)."""
    ndof = F.size
    dens = simp(x.ravel(), emin, emax, p)
    U = solve(assemble(KE, dens, edof, ndof), F, freedofs)
    ce = strain_energy(U, edof, KE)
    c = float(np.sum(dens * ce))
    dcx = -p * (emax - emin) * x.ravel() ** (p - 1) * ce
    return c, dcx, U


def gravity_compliance(x, t, ti, beta, KE, edof, freedofs, fe, emin=1e-9, emax=1.0, p=3.0):
    """`(c, dc/dx, dc/dt, U)` of one self-weight stage
    (tests/reference/compliance.py:51-110)."""
    nely, nelx = x.shape
    nel = nelx * nely
    m = time_mask(t, ti, beta).ravel()
    dm = time_mask_derivative(t, ti, beta).ravel()
    xr = x.ravel()
    xt = xr * m
    C = gravity_matrix(nelx, nely) * (fe * nel)
    ndof = 2 * (nelx + 1) * (nely + 1)
    F = np.zeros(ndof)
    F[1::2] = -(C @ xt)
    dens = simp(xt, emin, emax, p)
    U = solve(assemble(KE, dens, edof, ndof), F, freedofs)
    ce = strain_energy(U, edof, KE)
    c = float(np.sum(dens * ce))
    A = -p * (emax - emin) * xt ** (p - 1) * ce
    adj = -(C.T @ U[1::2])
    dcx = (A + 2 * adj) * m
    dct = (A + 2 * adj) * xr * dm
    return c, dcx, dct, U


def space_time(x, t, times, beta, KE, edof, F, freedofs, fe, emin=1e-9, emax=1.0, p=3.0):
    """Whole structure + every stage: `(c, dcx, [(cg, dcx_g, dct_g)], U rows)`."""
    c, dcx, U0 = whole_compliance(x, KE, edof, F, freedofs, emin, emax, p)
    stages, Us = [], [U0]
    for ti in times:
        cg, dx, dt, U = gravity_compliance(x, t, ti, beta, KE, edof, freedofs, fe, emin, emax, p)
        stages.append((cg, dx, dt))
        Us.append(U)
    return c, dcx, stages, np.stack(Us)

"""Pins the oracle itself (CPU only): closed forms, brute force on tiny meshes,
independent derivations and central finite differences, chosen so that a dropped
term, a wrong sign or index, or a transposed operand fails one of them."""

import math

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import fem_direct, mgcg_port

EMIN, EMAX, PENAL, NU = 1e-9, 1.0, 3.0, 0.3


def _pkg():
    import sttopt_amd

    return sttopt_amd


# ----------------------------------------------------------- element data ----


def test_element_tables_equal_quadrature_derivation():
    """Closed-form tables (oracle) == 2x2 Gauss quadrature (product): two derivations."""
    fem = _pkg().fem
    for nu in (0.0, 0.3, 0.45):
        np.testing.assert_allclose(fem_direct.plane_stress_KE(nu), fem.plane_stress_KE(nu), atol=1e-14)
    np.testing.assert_allclose(fem_direct.diffusion_KE(), fem.diffusion_KE(), atol=1e-15)


def test_plane_stress_KE_rigid_modes_and_uniform_strain_energy():
    KE = fem_direct.plane_stress_KE(NU)
    # corners BL, BR, TR, TL on the unit square (y up)
    xy = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    tx = np.tile([1.0, 0.0], 4)
    ty = np.tile([0.0, 1.0], 4)
    rot = np.ravel(np.stack([-xy[:, 1], xy[:, 0]], 1))
    for mode in (tx, ty, rot):
        assert np.abs(KE @ mode).max() < 1e-14
    w = np.linalg.eigvalsh(KE)
    assert np.sum(np.abs(w) < 1e-12) == 3 and np.all(w[3:] > 1e-3)
    # u = (x, 0): eps_xx = 1, energy u^T K u = E / (1 - nu^2) (area 1)
    u = np.ravel(np.stack([xy[:, 0], np.zeros(4)], 1))
    assert abs(u @ KE @ u - 1 / (1 - NU**2)) < 1e-14
    # pure shear u = (y, x): gamma = 2, energy = G gamma^2 = 2 (1 - nu) ... / (1 - nu^2)
    u = np.ravel(np.stack([xy[:, 1], xy[:, 0]], 1))
    G = 1 / (2 * (1 + NU))
    assert abs(u @ KE @ u - G * 4) < 1e-14


def test_diffusion_KE_constant_and_linear_fields():
    KE = fem_direct.diffusion_KE()
    assert np.abs(KE.sum(1)).max() < 1e-15
    x = np.array([0, 1, 1, 0.0])
    assert abs(x @ KE @ x - 1.0) < 1e-15  # |grad u|^2 = 1 over a unit area


def test_dof_maps_agree():
    fem = _pkg().fem
    for nelx, nely, d in [(3, 2, 2), (5, 4, 1), (7, 3, 2)]:
        np.testing.assert_array_equal(fem_direct.element_dofs(nelx, nely, d),
                                      fem.element_dof_map(nelx, nely, d))


def test_assembly_equals_dense_brute_force():
    rng = np.random.default_rng(0)
    nelx, nely = 3, 2
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    ndof = 2 * (nelx + 1) * (nely + 1)
    dens = rng.uniform(0.1, 1, nelx * nely)
    dense = np.zeros((ndof, ndof))
    for e in range(nelx * nely):
        for a in range(8):
            for b in range(8):
                dense[edof[e, a], edof[e, b]] += dens[e] * KE[a, b]
    np.testing.assert_allclose(fem_direct.assemble(KE, dens, edof, ndof).toarray(), dense, atol=1e-15)


# ----------------------------------------------------------- closed forms ----


def test_uniaxial_tension_patch_test():
    """Bar under uniform end traction: u_x = sigma x, u_y = -nu sigma y exactly."""
    fem = _pkg().fem
    nelx, nely, P = 8, 3, 2.0
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    nodes = fem.node_grid(nelx, nely)
    ndof = 2 * nodes.size
    F = np.zeros(ndof)
    right_up = nodes[::-1, -1]
    F[2 * right_up] = P * fem.edge_load_shares(nely + 1, nely)
    fixed = np.concatenate([2 * nodes[:, 0], [2 * nodes[-1, 0] + 1]])
    free = np.setdiff1d(np.arange(ndof), fixed)
    U = fem_direct.solve(fem_direct.assemble(KE, np.ones(nelx * nely), edof, ndof), F, free)
    sigma = P / nely
    r, c = np.divmod(np.arange(nodes.size), nelx + 1)
    y = nely - r
    np.testing.assert_allclose(U[0::2], sigma * c, atol=1e-12)
    np.testing.assert_allclose(U[1::2], -NU * sigma * y, atol=1e-12)
    # compliance = P * u_tip = P * sigma * L
    assert abs(F @ U - P * sigma * nelx) < 1e-10


def test_self_weight_matrix_totals_and_reactions():
    nelx, nely = 6, 4
    C = fem_direct.gravity_matrix(nelx, nely)
    x = np.random.default_rng(1).uniform(0.1, 1, nelx * nely)
    assert abs((C @ x).sum() - x.sum() / (nelx * nely)) < 1e-15
    np.testing.assert_allclose(C.T @ np.ones(C.shape[0]), np.full(nelx * nely, 1 / (nelx * nely)))
    fem = _pkg().fem
    np.testing.assert_allclose(fem.gravity_load_matrix(nelx, nely).toarray(), C.toarray())
    # hanging column, bottom clamped, unit density: reactions balance the unit weight
    F, free = fem.load_case("bottom_clamped", nelx, nely)
    Fy = np.zeros_like(F)
    Fy[1::2] = -(C @ np.ones(nelx * nely))
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    K = fem_direct.assemble(KE, np.ones(nelx * nely), edof, F.size)
    U = fem_direct.solve(K, Fy, free)
    reaction = K @ U - Fy
    fixed = np.setdiff1d(np.arange(F.size), free)
    assert abs(reaction[fixed][1::2].sum() - 1.0) < 1e-10


def test_mbb_load_case_dofs():
    fem = _pkg().fem
    nelx, nely = 120, 40
    F, free = fem.load_case("mbb", nelx, nely)
    assert F.size == 9922 and F.sum() == -1.0 and F[1] == -1.0
    nodes = fem.node_grid(nelx, nely)
    fixed = np.setdiff1d(np.arange(F.size), free)
    want = np.sort(np.concatenate([2 * nodes[:, 0], [2 * nodes[-1, -1] + 1]]))
    np.testing.assert_array_equal(fixed, want)


# ------------------------------------------------------ sensitivities (FD) ----


def _fd(f, x, idx, h):
    xp, xm = x.copy(), x.copy()
    xp.flat[idx] += h
    xm.flat[idx] -= h
    return (f(xp) - f(xm)) / (2 * h)


def test_time_mask_derivative_matches_fd():
    t = np.linspace(0.01, 0.99, 23)
    for ti, beta in [(0.3, 16.0), (0.7, 4.0)]:
        d = fem_direct.time_mask_derivative(t, ti, beta)
        h = 1e-6
        fd = (fem_direct.time_mask(t + h, ti, beta) - fem_direct.time_mask(t - h, ti, beta)) / (2 * h)
        np.testing.assert_allclose(d, fd, rtol=1e-6, atol=1e-8)


def test_hand_derived_sensitivities_match_fd():
    """Whole-structure dc/dx and the stage dc/dx, dc/dt (incl. the doubled
    gravity-load adjoint term) against central differences of the direct solve."""
    fem = _pkg().fem
    nelx, nely = 8, 5
    rng = np.random.default_rng(3)
    x = rng.uniform(0.3, 0.7, (nely, nelx))
    # stage well into the build (every element partly active): a well-conditioned
    # system, so the central difference is not dominated by cancellation
    t = rng.uniform(0.0, 0.6, (nely, nelx))
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    F, free = fem.load_case("cantilever", nelx, nely)
    _, dcx, _ = fem_direct.whole_compliance(x, KE, edof, F, free)
    Fb, freeb = fem.load_case("bottom_clamped", nelx, nely)
    ti, beta, fe = 0.7, 4.0, 1.0 / (nelx * nely)
    _, gx, gt, _ = fem_direct.gravity_compliance(x, t, ti, beta, KE, edof, freeb, fe)
    for e in rng.choice(nelx * nely, 6, replace=False):
        fd = _fd(lambda xx: fem_direct.whole_compliance(xx, KE, edof, F, free)[0], x, e, 1e-6)
        assert abs(fd - dcx[e]) <= 1e-5 * abs(dcx[e])
        fd = _fd(lambda xx: fem_direct.gravity_compliance(xx, t, ti, beta, KE, edof, freeb, fe)[0], x, e, 1e-6)
        assert abs(fd - gx[e]) <= 1e-5 * max(abs(gx[e]), 1e-3 * np.abs(gx).max())
        fd = _fd(lambda tt: fem_direct.gravity_compliance(x, tt, ti, beta, KE, edof, freeb, fe)[0], t, e, 1e-6)
        assert abs(fd - gt[e]) <= 1e-5 * max(abs(gt[e]), 1e-3 * np.abs(gt).max())


# ------------------------------------------------------------ the MG port ----


def _port_setup(nelx, nely, x):
    fem = _pkg().fem
    F, free = fem.load_case("cantilever", nelx, nely)
    ndof = F.size
    mask = np.zeros(ndof, bool)
    mask[free] = True
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    dens = fem_direct.simp(x, EMIN, EMAX, PENAL)
    return F, free, mask, KE, edof, dens, ndof


def test_port_galerkin_adjoint_transfers_and_spd_vcycle():
    from conftest import binary_design

    nelx, nely = 18, 12
    F, free, mask, KE, edof, dens, ndof = _port_setup(nelx, nely, binary_design(nelx, nely))
    T = torch.as_tensor
    levels = mgcg_port.build_hierarchy(T(dens), T(edof), T(KE), T(mask), nelx, nely, max_coarse_elements=4)
    assert len(levels) >= 3
    rng = np.random.default_rng(0)
    for f, c in zip(levels, levels[1:]):
        v = T(rng.standard_normal(c.ndof)) * c.mask
        Pv = mgcg_port.on_node_grid(v, c.nelx + 1, c.nely + 1, f.kx, f.ky, mgcg_port.interp_axis) * f.mask
        RAPv = mgcg_port.on_node_grid(f(Pv), f.nelx + 1, f.nely + 1, f.kx, f.ky, mgcg_port.restrict_axis) * c.mask
        np.testing.assert_allclose(c(v).numpy(), RAPv.numpy(), rtol=1e-10, atol=1e-12 * float(RAPv.abs().max()))
    M = mgcg_port.VCycle(mgcg_port.build_hierarchy(T(dens), T(edof), T(KE), T(mask), nelx, nely,
                                                   max_coarse_elements=24))
    u = T(rng.standard_normal(ndof)) * T(mask)
    w = T(rng.standard_normal(ndof)) * T(mask)
    a, b = float((M(u) * w).sum()), float((u * M(w)).sum())
    assert abs(a - b) <= 1e-10 * abs(a) and float((u * M(u)).sum()) > 0


def test_port_mgcg_and_jacobi_match_direct():
    nelx, nely = 18, 12
    x = np.random.default_rng(2).uniform(0.05, 1, nelx * nely)
    F, free, mask, KE, edof, dens, ndof = _port_setup(nelx, nely, x)
    want = fem_direct.solve(fem_direct.assemble(KE, dens, edof, ndof), F, free)
    T = torch.as_tensor
    U, n, _ = mgcg_port.mg_solve(T(F), T(dens), T(edof), T(KE), T(mask), nelx, nely, rtol=1e-11,
                                 max_coarse_elements=24)
    np.testing.assert_allclose(U.numpy(), want, atol=1e-8 * np.abs(want).max())
    diag = mgcg_port.diagonal(T(dens), T(edof), T(KE), ndof)
    diag = torch.where(T(mask), diag, torch.ones(()))
    Uj, nj = mgcg_port.pcg(lambda v: mgcg_port.operator(v, T(dens), T(edof), T(KE), ndof, T(mask)),
                           T(F) * T(mask), lambda r: r / diag, rtol=1e-11, max_iter=20000)
    np.testing.assert_allclose(Uj.numpy(), want, atol=1e-7 * np.abs(want).max())
    assert n < nj / 5


def test_port_adjoint_is_twice_U_for_compliance():
    nelx, nely = 12, 8
    x = np.random.default_rng(4).uniform(0.2, 0.9, nelx * nely)
    F, free, mask, KE, edof, dens, ndof = _port_setup(nelx, nely, x)
    T = torch.as_tensor
    U, _, levels = mgcg_port.mg_solve(T(F), T(dens), T(edof), T(KE), T(mask), nelx, nely, rtol=1e-10,
                                      max_coarse_elements=24)
    g = 2 * mgcg_port.matvec(U, T(dens), T(edof), T(KE), ndof)
    lam, grad, n = mgcg_port.adjoint(levels, U, T(F), g, T(edof), T(KE), rtol=1e-8)
    assert n == 0
    np.testing.assert_allclose(lam.numpy(), 2 * U.numpy(), rtol=1e-7, atol=1e-12)
    ce = fem_direct.strain_energy(U.numpy(), edof, KE)
    np.testing.assert_allclose(grad.numpy(), -2 * ce, rtol=1e-6, atol=1e-14)


def test_port_retirement_matches_single_solves():
    from conftest import binary_design

    nelx, nely = 18, 12
    xs = [binary_design(nelx, nely), np.full(nelx * nely, 0.4), np.full(nelx * nely, 0.9)]
    F, free, mask, KE, edof, _, ndof = _port_setup(nelx, nely, xs[0])
    T = torch.as_tensor
    dens = T(np.stack([fem_direct.simp(x, EMIN, EMAX, PENAL) for x in xs]))
    Ub, nb, _ = mgcg_port.mg_solve(T(F).expand(3, -1), dens, T(edof), T(KE), T(mask), nelx, nely,
                                   rtol=1e-11, max_coarse_elements=24)
    ns = []
    for k in range(3):
        U, n, _ = mgcg_port.mg_solve(T(F), dens[k], T(edof), T(KE), T(mask), nelx, nely, rtol=1e-11,
                                     max_coarse_elements=24)
        ns.append(n)
        np.testing.assert_allclose(Ub[k].numpy(), U.numpy(), atol=1e-6 * float(U.abs().max()))
    assert nb == max(ns) and min(ns) < max(ns)


def test_sparse_gravity_scaling_in_stage_oracle():
    """gravity_compliance scales C to the requested fe: doubling fe quadruples c."""
    fem = _pkg().fem
    nelx, nely = 6, 4
    rng = np.random.default_rng(5)
    x = rng.uniform(0.3, 0.7, (nely, nelx))
    t = rng.uniform(0, 1, (nely, nelx))
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    _, free = fem.load_case("bottom_clamped", nelx, nely)
    fe = 1 / (nelx * nely)
    c1 = fem_direct.gravity_compliance(x, t, 0.6, 8.0, KE, edof, free, fe)[0]
    c2 = fem_direct.gravity_compliance(x, t, 0.6, 8.0, KE, edof, free, 2 * fe)[0]
    assert abs(c2 / c1 - 4.0) < 1e-10
    assert isinstance(fem_direct.gravity_matrix(nelx, nely), sp.csr_matrix)


@pytest.mark.parametrize("nelx, nely", [(120, 40)])
def test_c1_mbb_oracle_solution(nelx, nely):
    """The plain-method configuration (120x40 MBB, 9,922 dofs) through the oracle."""
    wl = _pkg().workloads
    p = wl.mbb(nelx, nely, device="cpu")
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    c, dcx, U = fem_direct.whole_compliance(p.xPhys.numpy(), KE, edof, p.F.numpy(), p.freedofs.numpy())
    assert p.ndof == 9922 and c > 0 and np.all(dcx < 0)
    # compliance equals F . U (work of the external load)
    assert abs(c - float(p.F.numpy() @ U)) <= 1e-9 * c


# ------------------------------------------------ scalar-mesh gradient fields ----


def test_heat_shape_gradients_are_the_shape_function_derivatives():
    """B against central differences of the four shape functions at the Gauss points,
    partition of unity (rows sum to 0), and the eta-major Gauss-point order."""
    from oracle import heat_fields

    B = heat_fields.shape_gradients()
    h = 1e-6
    for g, (xi, eta) in enumerate(heat_fields.gauss_points()):
        dx = (heat_fields.shape_functions(xi + h, eta) - heat_fields.shape_functions(xi - h, eta)) / (2 * h)
        dy = (heat_fields.shape_functions(xi, eta + h) - heat_fields.shape_functions(xi, eta - h)) / (2 * h)
        np.testing.assert_allclose(B[g, 0], dx, atol=1e-9)
        np.testing.assert_allclose(B[g, 1], dy, atol=1e-9)
    np.testing.assert_allclose(B.sum(-1), 0.0, atol=1e-15)
    pts = heat_fields.gauss_points()
    assert pts[0][0] < 0.5 and pts[0][1] < 0.5 and pts[1][0] > 0.5 and pts[2][1] > 0.5


@pytest.mark.parametrize("nelx, nely", [(3, 2), (7, 5)])
def test_heat_gauss_gradients_of_coordinate_and_bilinear_fields(nelx, nely):
    """x (column) and y (up: nely - row) have gradients (1, 0) and (0, 1); x y has
    (y_g, x_g) at each Gauss point's global coordinates (exact for bilinear fields)."""
    from oracle import heat_fields

    r, c = np.divmod(np.arange((nelx + 1) * (nely + 1)), nelx + 1)
    x, y = c.astype(float), (nely - r).astype(float)
    np.testing.assert_allclose(heat_fields.gauss_gradients(x, nelx, nely), np.broadcast_to([1.0, 0.0], (nelx * nely, 4, 2)), atol=1e-14)
    np.testing.assert_allclose(heat_fields.gauss_gradients(y, nelx, nely), np.broadcast_to([0.0, 1.0], (nelx * nely, 4, 2)), atol=1e-14)
    G = heat_fields.gauss_gradients(x * y, nelx, nely)
    er, ec = np.divmod(np.arange(nelx * nely), nelx)
    for g, (xi, eta) in enumerate(heat_fields.gauss_points()):
        xg = ec + xi
        yg = (nely - er - 1) + eta  # element row er spans y in [nely - er - 1, nely - er]
        np.testing.assert_allclose(G[:, g, 0], yg, atol=1e-13)
        np.testing.assert_allclose(G[:, g, 1], xg, atol=1e-13)


def test_heat_divergence_load_of_a_constant_field_is_the_boundary_flux():
    """w = 1, X = (1, 0): F = integral of dN_i/dx = the outward x-flux through the
    boundary: -1/2 per vertical boundary edge at the left nodes, +1/2 at the right,
    0 inside; X = (0, 1) the same through the bottom (-) and top (+) edges."""
    from oracle import heat_fields

    nelx, nely = 6, 4
    nel = nelx * nely
    w = np.ones(nel)
    F = heat_fields.divergence_load(w, np.broadcast_to([1.0, 0.0], (nel, 4, 2)), nelx, nely)
    want = np.zeros((nely + 1, nelx + 1))
    edges = np.full(nely + 1, 2.0)
    edges[[0, -1]] = 1.0
    want[:, 0] = -0.5 * edges
    want[:, -1] = 0.5 * edges
    np.testing.assert_allclose(F.reshape(nely + 1, nelx + 1), want, atol=1e-15)
    F = heat_fields.divergence_load(w, np.broadcast_to([0.0, 1.0], (nel, 4, 2)), nelx, nely)
    want = np.zeros((nely + 1, nelx + 1))
    edges = np.full(nelx + 1, 2.0)
    edges[[0, -1]] = 1.0
    want[0, :] = 0.5 * edges  # top row (y up)
    want[-1, :] = -0.5 * edges
    np.testing.assert_allclose(F.reshape(nely + 1, nelx + 1), want, atol=1e-15)


def test_heat_divergence_load_is_the_adjoint_of_the_gradients():
    """F(w, X) . phi = 1/4 sum_e w_e X_e : G(phi)_e for random fields: the element scatter
    and the element gather use the same node map; and the direction field is a unit
    field parallel to the gradient (0 at a flat element)."""
    from oracle import heat_fields

    rng = np.random.default_rng(0)
    nelx, nely = 5, 3
    nel, nn = nelx * nely, (nelx + 1) * (nely + 1)
    w = rng.uniform(0.1, 1, nel)
    X = rng.standard_normal((nel, 4, 2))
    phi = rng.standard_normal(nn)
    lhs = heat_fields.divergence_load(w, X, nelx, nely) @ phi
    rhs = 0.25 * np.einsum("e,egd,egd->", w, X, heat_fields.gauss_gradients(phi, nelx, nely))
    assert abs(lhs - rhs) < 1e-12 * abs(rhs)
    s = rng.standard_normal(nn)
    Y = heat_fields.direction_field(s, nelx, nely)
    G = heat_fields.gauss_gradients(s, nelx, nely)
    np.testing.assert_allclose(np.linalg.norm(Y, axis=-1), 1.0, atol=1e-14)
    np.testing.assert_allclose(Y[..., 0] * G[..., 1] - Y[..., 1] * G[..., 0], 0.0, atol=1e-14)
    assert np.all((Y * G).sum(-1) > 0)
    assert np.all(heat_fields.direction_field(np.ones(nn), nelx, nely) == 0.0)


# ------------------------------------------------ density filter, projection ----


def test_filter_oracle_closed_forms():
    from oracle import filters_ref

    nelx, nely = 7, 5
    H, Hs = filters_ref.filter_matrix(nelx, nely, 1.5)
    assert np.array_equal(H, H.T)
    np.testing.assert_allclose(Hs, H.sum(1))
    e = 2 * nelx + 3  # interior element (3, 2)
    row = H[e].reshape(nely, nelx)
    want = np.zeros((nely, nelx))
    want[1:4, 2:5] = 1.5 - np.sqrt(2)
    want[2, 2] = want[2, 4] = want[1, 3] = want[3, 3] = 0.5
    want[2, 3] = 1.5
    np.testing.assert_allclose(row, want, atol=1e-15)
    # corner (0, 0): itself, two edge neighbours, one diagonal
    assert abs(Hs[0] - (1.5 + 2 * 0.5 + (1.5 - np.sqrt(2)))) < 1e-15
    x = np.random.default_rng(0).uniform(size=(nely, nelx))
    np.testing.assert_allclose(filters_ref.apply(np.full((nely, nelx), 0.3), H, Hs), 0.3, atol=1e-15)
    for rmin in (0.5, 1.0):
        H1, Hs1 = filters_ref.filter_matrix(nelx, nely, rmin)
        np.testing.assert_allclose(filters_ref.apply(x, H1, Hs1), x, atol=1e-15)


def test_heaviside_oracle_closed_forms():
    from oracle import filters_ref

    x = np.linspace(-0.2, 1.2, 57)
    for beta, eta in ((1.0, 0.5), (8.0, 0.3), (32.0, 0.6)):
        np.testing.assert_allclose(filters_ref.heaviside(np.array([0.0, 1.0]), beta, eta), [0.0, 1.0], atol=1e-15)
        h = 1e-6
        fd = (filters_ref.heaviside(x + h, beta, eta) - filters_ref.heaviside(x - h, beta, eta)) / (2 * h)
        np.testing.assert_allclose(filters_ref.heaviside_derivative(x, beta, eta), fd, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(filters_ref.heaviside(x, 1e-4, 0.5), x, atol=1e-8)
    assert filters_ref.heaviside(np.array([0.29, 0.3, 0.31]), math.inf, 0.3).tolist() == [0.0, 0.5, 1.0]


# ------------------------------------------------------------------- MMA ----


def _mma_problem(m, n, seed):
    rng = np.random.default_rng(seed)
    return dict(low=np.full(n, -1.0), upp=np.full(n, 1.0), alfa=np.full(n, -0.9),
                beta=np.full(n, 0.9), p0=rng.uniform(0.5, 2.0, n), q0=rng.uniform(0.5, 2.0, n),
                P=rng.uniform(0.1, 1.0, (m, n)), Q=rng.uniform(0.1, 1.0, (m, n)), a0=1.0,
                a=rng.uniform(0.1, 0.3, m), c=np.full(m, 10.0), d=rng.uniform(0.25, 1.0, m))


@pytest.mark.parametrize("m, n", [(3, 8), (3, 2), (5, 40)])
def test_mma_subsolv_oracle_satisfies_kkt(m, n):
    """The barrier subproblem's KKT conditions at the returned point (epsi = 1e-6, the
    last barrier level), with binding constraints (b below each row's value at x = 0)."""
    from oracle import mma_ref

    pr = _mma_problem(m, n, m * 10 + n)
    pr["b"] = 0.8 * (pr["P"].sum(1) + pr["Q"].sum(1))
    sol = mma_ref.subsolv(m, n, 1e-7, pr["low"], pr["upp"], pr["alfa"], pr["beta"], pr["p0"],
                          pr["q0"], pr["P"], pr["Q"], pr["a0"], pr["a"], pr["b"], pr["c"], pr["d"])
    stat, comp = mma_ref.kkt_residuals(sol, **{k: pr[k] for k in ("low", "upp", "alfa", "beta", "p0", "q0",
                                                                  "P", "Q", "a0", "a", "b", "c", "d")})
    assert stat < 1e-5
    for prod in comp:
        assert np.all(np.abs(prod - 1e-6) < 1e-5)
    x, lam = sol[0], sol[3]
    assert np.all(x > pr["alfa"]) and np.all(x < pr["beta"]) and lam.max() > 1e-3


def test_mma_oracle_asymptote_update_golden_case():
    """The reference's hand-picked asymptote case (tests/test_mma.py:62-117): oscillating,
    monotone and stationary variables and each of the four clamps."""
    from oracle import mma_ref

    n, m = 6, 1
    xval = np.full(n, 5.0)
    res = mma_ref.mmasub(m, n, 3, xval, np.zeros(n), np.full(n, 10.0),
                         np.array([4.0, 4.0, 5.0, 4.0, 4.0, 4.0]), np.array([3.0, 5.0, 2.0, 3.0, 3.0, 4.0]),
                         np.zeros(n), np.zeros(m), np.zeros((m, n)),
                         np.array([2.0, 2.0, 2.0, -1000.0, 2.0, 4.8]),
                         np.array([8.0, 8.0, 8.0, 8.0, 1000.0, 4.0]), 1.0, np.zeros(m),
                         np.full(m, 2500.0), np.zeros(m))
    np.testing.assert_allclose(res[-2], [2.6, 3.6, 2.0, -95.0, 2.6, 4.9], atol=1e-12)
    np.testing.assert_allclose(res[-1], [9.8, 7.8, 8.0, 9.8, 105.0, 5.1], atol=1e-12)


def _mma_toy(mmasub, beam: bool):
    n, m = (5, 1) if beam else (2, 1)
    coef = np.array([61.0, 37.0, 19.0, 7.0, 1.0])
    if beam:
        obj = lambda x: (0.0624 * x.sum(), np.full(n, 0.0624))  # noqa: E731
        con = lambda x: (np.array([np.sum(coef / x**3) - 1.0]), (-3 * coef / x**4)[None, :])  # noqa: E731
        xmin, xmax, x, iters = np.full(n, 0.1), np.full(n, 10.0), np.full(n, 5.0), 20
    else:
        obj = lambda x: (np.sum((x - 4) ** 2), 2 * (x - 4))  # noqa: E731
        con = lambda x: (np.array([x.sum() - 2.0]), np.ones((1, n)))  # noqa: E731
        xmin, xmax, x, iters = np.full(n, -10.0), np.full(n, 10.0), np.zeros(n), 10
    xold1, xold2 = x.copy(), x.copy()
    low, upp = np.zeros(n), np.zeros(n)
    for it in range(iters):
        _, g0 = obj(x)
        fv, g = con(x)
        res = mmasub(m, n, it, x, xmin, xmax, xold1, xold2, g0, fv, g, low, upp, 1.0,
                     np.zeros(m), np.array([1000.0]), np.zeros(m))
        xold2, xold1 = xold1, x
        x, low, upp = np.asarray(res[0]), np.asarray(res[-2]), np.asarray(res[-1])
    return x


def test_mma_oracle_converges_to_closed_form_optima():
    """Svanberg's cantilever beam (x* proportional to coef^(1/4), the constraint active)
    and min |x - (4, 4)|^2 s.t. x1 + x2 <= 2 (x* = (1, 1))."""
    from oracle import mma_ref

    coef = np.array([61.0, 37.0, 19.0, 7.0, 1.0])
    root4 = coef**0.25
    np.testing.assert_allclose(_mma_toy(mma_ref.mmasub, True), np.sum(root4) ** (1 / 3) * root4, rtol=1e-4)
    np.testing.assert_allclose(_mma_toy(mma_ref.mmasub, False), [1.0, 1.0], atol=1e-6)


# ------------------------------- the V-cycle as explicit dense matrices ----


def _hat_interpolation(nc: int, k: int) -> np.ndarray:
    """1-D prolongation on nodes: fine node i sits at coarse coordinate i / k and takes
    the coarse nodal hat functions' values there, max(0, 1 - |i / k - j|)."""
    fine = np.arange(nc * k + 1)[:, None] / k
    return np.maximum(0.0, 1.0 - np.abs(fine - np.arange(nc + 1)[None, :]))


def _dense_vcycle_operator(K0, mask0, nelx, nely, d, omega, n_smooth, gamma, max_coarse):
    """The V-cycle preconditioner as one dense matrix, from global matrices only.

    Level l: unprojected operator A_l (A_0 = the assembled K, A_{l+1} = P_l^T A_l P_l),
    projected operator M_l A_l M_l, Jacobi diagonal D_l = diag(A_l) with 1 at fixed dofs,
    prolongation P_l = kron(P_y, P_x, I_d) of the nodal hat functions, coarse mask =
    the fine mask at the nodes (r ky, c kx), coarsest solve (M A M + I - M)^-1 M.
    Smoothing x <- x + omega D^-1 (b - M A M x), n_smooth times, then x <- M x.  The
    right-hand side is projected on entry (z = B M_0 r)."""
    levels = []
    A, mask, nx, ny = K0, mask0.astype(float), nelx, nely
    while True:
        levels.append((A, mask, nx, ny))
        if nx * ny <= max_coarse:
            break
        kx = 2 if nx % 2 == 0 else (3 if nx % 3 == 0 else 1)
        ky = 2 if ny % 2 == 0 else (3 if ny % 3 == 0 else 1)
        if (kx, ky) == (1, 1):
            break
        cx, cy = nx // kx, ny // ky
        P = np.kron(np.kron(_hat_interpolation(cy, ky), _hat_interpolation(cx, kx)), np.eye(d))
        grid = mask.reshape(ny + 1, nx + 1, d)
        levels[-1] = levels[-1] + (P,)
        A, mask, nx, ny = P.T @ A @ P, grid[::ky, ::kx].reshape(-1), cx, cy

    def B(i):
        A, m = levels[i][0], np.diag(levels[i][1])
        n = A.shape[0]
        Ah = m @ A @ m
        if i == len(levels) - 1:
            return m @ np.linalg.inv(Ah + np.eye(n) - m) @ m
        D = np.where(levels[i][1] > 0, np.diag(A), 1.0)
        W = omega * np.diag(1.0 / D)
        J = np.eye(n) - W @ Ah
        P = levels[i][4]
        mc = np.diag(levels[i + 1][1])
        Ac = mc @ levels[i + 1][0] @ mc
        Bc = B(i + 1)
        # coarse correction with gamma cycles: E = sum of (I - Bc Ac)^j Bc over j < gamma
        E = np.zeros_like(Bc)
        Tc = np.eye(Bc.shape[0])
        for _ in range(gamma):
            E = E + Tc @ Bc
            Tc = (np.eye(Bc.shape[0]) - Bc @ Ac) @ Tc
        # pre-smoothing from 0: x = m S b, S = sum_{j < nu} J^j W
        S = sum(np.linalg.matrix_power(J, j) for j in range(n_smooth)) @ W
        X1 = m @ S
        R1 = np.eye(n) - Ah @ X1  # r = R1 b
        X2 = X1 + m @ P @ mc @ E @ mc @ P.T @ R1
        # post-smoothing: x <- J x + W b, nu times, then x <- m x
        Jn = np.linalg.matrix_power(J, n_smooth)
        return m @ (Jn @ X2 + S)

    return B(0) @ np.diag(levels[0][1]), len(levels)


@pytest.mark.parametrize(
    "nelx,nely,d,omega,n_smooth,gamma,max_coarse,reaction,nlev",
    [
        (8, 4, 2, 0.6, 2, 1, 8, False, 2),     # the reference's defaults, 2 levels
        (12, 9, 2, 0.6, 2, 1, 3, False, 3),    # factor 3 in y, 3 levels
        (12, 9, 2, 0.45, 3, 1, 3, False, 3),   # another omega and sweep count
        (8, 6, 2, 0.6, 1, 2, 3, False, 3),     # W-cycle (gamma 2), one sweep
        (12, 6, 1, 0.7, 2, 1, 2, True, 4),     # scalar mesh with a lumped reaction, 4 levels
    ],
)
def test_port_vcycle_equals_dense_matrix_vcycle(nelx, nely, d, omega, n_smooth, gamma,
                                                 max_coarse, reaction, nlev):
    """mgcg_port.VCycle (element Galerkin, node-grid transfers, index_add applies,
    Cholesky coarsest solve) == the same cycle written as products of global dense
    matrices, to 1e-12.  A wrong omega, sweep count, smoothing order, transfer weight or
    coarse mask changes the dense operator."""
    rng = np.random.default_rng(7)
    nel = nelx * nely
    nn = (nelx + 1) * (nely + 1)
    ndof = d * nn
    KE = fem_direct.plane_stress_KE(NU) if d == 2 else fem_direct.diffusion_KE()
    edof = fem_direct.element_dofs(nelx, nely, d)
    dens = rng.uniform(0.05, 1.0, nel)
    rea = rng.uniform(0.1, 2.0, nel) if reaction else None
    mask = np.ones(ndof, bool)
    grid = mask.reshape(nely + 1, nelx + 1, d)
    if d == 2:
        grid[:, 0, :] = False            # clamped left edge
        grid[1, 1, 1] = False            # one fixed dof at a node off every coarse grid
    else:
        grid[-1, :, 0] = False           # Dirichlet bottom edge
        grid[2, 3, 0] = False
    K0 = fem_direct.assemble(KE, dens, edof, ndof, rea).toarray()
    Bd, nlev_dense = _dense_vcycle_operator(K0, mask, nelx, nely, d, omega, n_smooth, gamma,
                                            max_coarse)
    T = torch.as_tensor
    levels = mgcg_port.build_hierarchy(T(dens), T(edof), T(KE), T(mask), nelx, nely,
                                       reaction=None if rea is None else T(rea),
                                       max_coarse_elements=max_coarse)
    assert len(levels) == nlev_dense == nlev
    M = mgcg_port.VCycle(levels, omega=omega, n_smooth=n_smooth, gamma=gamma)
    r = rng.standard_normal((3, ndof))
    z = M(T(r)).numpy()
    want = r @ Bd.T
    np.testing.assert_allclose(z, want, rtol=0, atol=1e-12 * np.abs(want).max())
    # the dense operator itself is the SPD preconditioner the outer CG needs
    free = np.flatnonzero(mask)
    Bf = Bd[np.ix_(free, free)]
    np.testing.assert_allclose(Bf, Bf.T, atol=1e-12 * np.abs(Bf).max())
    assert np.linalg.eigvalsh((Bf + Bf.T) / 2).min() > 0


def test_dense_vcycle_construction_is_exact_on_one_level():
    """The dense construction's own pin: with the coarsest level at the top, the cycle
    is the exact inverse on the free dofs (M A M + I - M)^-1 restricted by the mask."""
    nelx, nely = 4, 3
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    ndof = 2 * (nelx + 1) * (nely + 1)
    mask = np.ones(ndof, bool)
    mask[:2 * (nelx + 1)] = False
    K0 = fem_direct.assemble(KE, np.full(nelx * nely, 0.7), edof, ndof).toarray()
    Bd, nlev = _dense_vcycle_operator(K0, mask, nelx, nely, 2, 0.6, 2, 1, 100)
    assert nlev == 1
    free = np.flatnonzero(mask)
    np.testing.assert_allclose(Bd[np.ix_(free, free)], np.linalg.inv(K0[np.ix_(free, free)]),
                               rtol=1e-10, atol=1e-12)
    # and the hat prolongation reproduces linear fields exactly
    P = _hat_interpolation(3, 2)
    np.testing.assert_allclose(P @ (2.0 + 0.5 * np.arange(4)), 2.0 + 0.25 * np.arange(7))
    np.testing.assert_allclose(P.sum(1), 1.0)


# ---------------------------------------------------- the lifted direct solve ----


@pytest.mark.parametrize("nelx,nely", [(5, 3), (8, 6)])
def test_lifted_solve_reproduces_linear_fields(nelx, nely):
    """Dirichlet data of a linear field u = a + b x + c y on the whole boundary (no
    source): the bilinear elements hold it exactly, so the lifted solve returns it at
    every node; and 0 at the bottom, 1 at the top with free sides: u = height fraction.
    A sign slip in F - K g, or a lift that ignores g, fails both."""
    KE = fem_direct.diffusion_KE()
    edof = fem_direct.element_dofs(nelx, nely, 1)
    nn = (nelx + 1) * (nely + 1)
    rows, cols = np.divmod(np.arange(nn), nelx + 1)
    height = (nely - rows).astype(float)   # row 0 is the top
    K = fem_direct.assemble(KE, np.ones(nelx * nely), edof, nn)
    exact = 0.3 + 0.7 * cols - 1.1 * height
    boundary = (rows == 0) | (rows == nely) | (cols == 0) | (cols == nelx)
    u = fem_direct.lifted_solve(K, np.zeros(nn), np.where(boundary, exact, 0.0), boundary)
    np.testing.assert_allclose(u, exact, atol=1e-12)
    plates = (rows == 0) | (rows == nely)
    g = np.where(rows == 0, 1.0, 0.0)
    u = fem_direct.lifted_solve(K, np.zeros(nn), g, plates)
    np.testing.assert_allclose(u, height / nely, atol=1e-12)
    # with a reaction term and a source the free rows satisfy K u = F exactly
    rea = np.random.default_rng(1).uniform(0.1, 1.0, nelx * nely)
    Kr = fem_direct.assemble(KE, np.ones(nelx * nely), edof, nn, rea)
    F = np.random.default_rng(2).standard_normal(nn)
    u = fem_direct.lifted_solve(Kr, F, g, plates)
    np.testing.assert_allclose(u[plates], g[plates])
    np.testing.assert_allclose((Kr @ u - F)[~plates], 0.0, atol=1e-11)

"""Gauss-point gradient fields of the scalar mesh, the divergence load, the Poisson
distance and the heat time field on the device against the NumPy oracle
(oracle/heat_fields.py, fem_direct's direct solves): forward values, the hand-written
adjoints against central differences of the oracle, and the time-field maps end to end
(sttopt/virtual_heat.py:290-462)."""

import types

import numpy as np
import pytest
import torch

from oracle import fem_direct, heat_fields

pytestmark = pytest.mark.gpu


def _mesh(nelx, nely, dev, base="bottom_edge"):
    from sttopt_amd import virtual_heat

    return virtual_heat.ScalarMesh.build(nelx, nely, base, dev, torch.float64)


def _t(a, dev):
    return torch.as_tensor(np.asarray(a), dtype=torch.float64, device=dev)


@pytest.mark.parametrize("nelx, nely, rows", [(7, 5, 1), (64, 31, 3), (130, 66, 2)])
def test_gradient_fields_and_load_match_oracle(dev, nelx, nely, rows):
    mesh = _mesh(nelx, nely, dev)
    rng = np.random.default_rng(nelx + rows)
    nn, nel = mesh.ndof, nelx * nely
    s = rng.standard_normal((rows, nn))
    w = rng.uniform(0.1, 1, (rows, nel))
    X = rng.standard_normal((rows, nel, 4, 2))
    G = mesh.gauss_gradients(_t(s, dev)).cpu().numpy()
    Y = mesh.direction_field(_t(s, dev)).cpu().numpy()
    F = mesh.divergence_load(_t(w, dev), _t(X, dev)).cpu().numpy()
    assert G.shape == (rows, nel, 4, 2) and F.shape == (rows, nn)
    for r in range(rows):
        g0 = heat_fields.gauss_gradients(s[r], nelx, nely)
        np.testing.assert_allclose(G[r], g0, rtol=0, atol=1e-13 * np.abs(g0).max())
        np.testing.assert_allclose(Y[r], heat_fields.direction_field(s[r], nelx, nely), rtol=0, atol=1e-13)
        f0 = heat_fields.divergence_load(w[r], X[r], nelx, nely)
        np.testing.assert_allclose(F[r], f0, rtol=0, atol=1e-13 * np.abs(f0).max())
    # one unbatched field shared by the rows of w (broadcast)
    F1 = mesh.divergence_load(_t(w, dev), _t(X[0], dev)).cpu().numpy()
    np.testing.assert_allclose(F1[-1], heat_fields.divergence_load(w[-1], X[0], nelx, nely), atol=1e-12)


def test_gradient_field_adjoints_match_finite_differences(dev):
    """Autograd through the hand-written adjoints (sttopt_direction_field_backward,
    sttopt_gauss_dot, the load with w = 1 as the adjoint of the gradients) against
    central differences of the oracle's forward maps."""
    nelx, nely = 6, 4
    mesh = _mesh(nelx, nely, dev)
    rng = np.random.default_rng(2)
    nn, nel = mesh.ndof, nelx * nely
    s = rng.standard_normal(nn)
    w = rng.uniform(0.2, 1, nel)
    X = rng.standard_normal((nel, 4, 2))
    cY = rng.standard_normal((nel, 4, 2))
    cG = rng.standard_normal((nel, 4, 2))
    cF = rng.standard_normal(nn)

    def fd(f, x, h):
        g = np.zeros_like(x)
        for i in np.ndindex(x.shape):
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            g[i] = (f(xp) - f(xm)) / (2 * h)
        return g

    st = _t(s, dev).requires_grad_(True)
    ((mesh.direction_field(st) * _t(cY, dev)).sum() + (mesh.gauss_gradients(st) * _t(cG, dev)).sum()).backward()
    want = fd(lambda v: (heat_fields.direction_field(v, nelx, nely) * cY).sum()
              + (heat_fields.gauss_gradients(v, nelx, nely) * cG).sum(), s, 1e-6)
    np.testing.assert_allclose(st.grad.cpu().numpy(), want, rtol=1e-6, atol=1e-7)

    wt, Xt = _t(w, dev).requires_grad_(True), _t(X, dev).requires_grad_(True)
    (mesh.divergence_load(wt, Xt) * _t(cF, dev)).sum().backward()
    np.testing.assert_allclose(wt.grad.cpu().numpy(),
                               fd(lambda v: heat_fields.divergence_load(v, X, nelx, nely) @ cF, w, 1e-6),
                               rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(Xt.grad.cpu().numpy(),
                               fd(lambda v: heat_fields.divergence_load(w, v, nelx, nely) @ cF, X, 1e-6),
                               rtol=1e-7, atol=1e-9)


def test_poisson_distance_matches_direct_solve(dev):
    """phi = 0 on the plate, div(w grad phi) = div(w X(s)) against the assembled direct
    solve, on a field s that rises away from the plate with noise."""
    nelx, nely = 48, 30
    mesh = _mesh(nelx, nely, dev)
    rng = np.random.default_rng(9)
    r, _ = np.divmod(np.arange(mesh.ndof), nelx + 1)
    s = (nely - r) / nely + 0.05 * rng.standard_normal(mesh.ndof)
    w = rng.uniform(0.05, 1, nelx * nely)
    phi = mesh.poisson_distance(_t(s, dev), _t(w, dev), 1e-12).cpu().numpy()
    plate = mesh.plate.cpu().numpy()
    want = heat_fields.poisson_distance(s, w, plate, nelx, nely)
    assert np.all(phi[plate] == 0)
    assert np.abs(phi - want).max() <= 1e-8 * np.abs(want).max()


def _oracle_time_field(mesh_np, xPhys, mu, cfg, conduction_penal=1.0):
    """The heat time field composed from the oracle's direct solves."""
    nelx, nely, plate = mesh_np
    nel, nn = nelx * nely, (nelx + 1) * (nely + 1)
    edof = fem_direct.element_dofs(nelx, nely, 1)
    x = xPhys.ravel()
    density = np.maximum(x, 1e-6)
    chi = np.maximum(x**conduction_penal, 1e-6) * cfg.chi_contrast ** (mu.ravel() - 0.5)
    alpha = cfg.drain_beta / (nelx * nely)
    K = fem_direct.assemble(fem_direct.diffusion_KE(), chi, edof, nn, reaction=np.full(nel, alpha))
    T = fem_direct.lifted_solve(K, np.zeros(nn), plate.astype(float), plate)
    part = x >= x.max() / 2
    if cfg.time_map == "one_minus":
        t = 1 - T
    elif cfg.time_map == "neg_log":
        floor = 1e-3 * T[edof].mean(-1)[part].min()
        t = -np.log(np.maximum(T, 0) + floor)
    else:
        t = heat_fields.poisson_distance(-T, density, plate, nelx, nely)
    te = t[edof].mean(-1)
    return (te / te[part].max()).reshape(nely, nelx), T, t


@pytest.mark.parametrize("time_map", ["one_minus", "neg_log", "poisson"])
def test_heat_time_field_maps_match_oracle(dev, time_map):
    from sttopt_amd import virtual_heat

    nelx, nely = 40, 24
    mesh = _mesh(nelx, nely, dev)
    rng = np.random.default_rng(3)
    xPhys = rng.uniform(0.1, 1, (nely, nelx))
    mu = rng.uniform(0.3, 0.7, (nely, nelx))
    cfg = types.SimpleNamespace(chi_contrast=10.0, drain_beta=4.0, time_map=time_map,
                                heat_cg_rtol=1e-12, poisson_cg_rtol=1e-12)
    sol = virtual_heat.heat_time_field(cfg, mesh, _t(xPhys, dev), _t(mu, dev))
    want, T, t = _oracle_time_field((nelx, nely, mesh.plate.cpu().numpy()), xPhys, mu, cfg)
    assert (sol.poisson is not None) == (time_map == "poisson")
    np.testing.assert_allclose(sol.primary.cpu().numpy(), T, atol=1e-9)
    np.testing.assert_allclose(sol.t.detach().cpu().numpy(), t, atol=1e-8 * np.abs(t).max())
    np.testing.assert_allclose(sol.tPhys.detach().cpu().numpy(), want, atol=1e-8)


def test_poisson_time_field_gradient_matches_finite_differences(dev):
    """The real (non-self-adjoint) adjoint chain of the Poisson map -- heat solve,
    direction field, divergence load, Poisson solve -- against central differences of
    the oracle's composition along seeded directions in mu and xPhys."""
    from sttopt_amd import virtual_heat

    nelx, nely = 20, 12
    mesh = _mesh(nelx, nely, dev)
    plate = mesh.plate.cpu().numpy()
    rng = np.random.default_rng(5)
    xPhys = rng.uniform(0.2, 1, (nely, nelx))
    mu = rng.uniform(0.3, 0.7, (nely, nelx))
    c = rng.standard_normal((nely, nelx))
    cfg = types.SimpleNamespace(chi_contrast=10.0, drain_beta=4.0, time_map="poisson",
                                heat_cg_rtol=1e-13, poisson_cg_rtol=1e-13)
    xt, mt = _t(xPhys, dev).requires_grad_(True), _t(mu, dev).requires_grad_(True)
    sol = virtual_heat.heat_time_field(cfg, mesh, xt, mt)
    # the normalization scale is a constant (detached): differentiate t_e directly
    (sol.tPhys * _t(c, dev)).sum().backward()

    scale_x = _oracle_time_field((nelx, nely, plate), xPhys, mu, cfg)
    h = 1e-6
    for k in range(3):
        dmu = rng.standard_normal(mu.shape)
        dx = rng.standard_normal(xPhys.shape) * 0.1
        # hold the normalization at its base value: both sides differentiate t_e / s0
        te0 = scale_x[2][fem_direct.element_dofs(nelx, nely, 1)].mean(-1)
        s0 = te0[(xPhys.ravel() >= xPhys.max() / 2)].max()

        def Lfixed(xv, mv):
            _, _, t = _oracle_time_field((nelx, nely, plate), xv, mv, cfg)
            te = t[fem_direct.element_dofs(nelx, nely, 1)].mean(-1).reshape(nely, nelx)
            return ((te / s0) * c).sum()

        fd = (Lfixed(xPhys + h * dx, mu + h * dmu) - Lfixed(xPhys - h * dx, mu - h * dmu)) / (2 * h)
        ad = float((xt.grad.cpu().numpy() * dx).sum() + (mt.grad.cpu().numpy() * dmu).sum())
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-8), (k, ad, fd)


# ------------------------------------------------------- Laplace time field ----


def _oracle_laplace(nelx, nely, active, plate, arc, xPhys, mu, b, cfg, e=0.0):
    """The Laplace time field composed from the oracle's direct solves: chi over the
    domain (1 outside), t = 0 on the plate, t = b on the wall arc."""
    nn = (nelx + 1) * (nely + 1)
    edof = fem_direct.element_dofs(nelx, nely, 1)
    x = xPhys.ravel()
    chi = cfg.chi_contrast ** (mu.ravel() - 0.5)
    if e != 0:
        chi = chi * np.maximum(x**e, cfg.chi_density_floor)
    chi = np.where(active.ravel(), chi, 1.0)
    K = fem_direct.assemble(fem_direct.diffusion_KE(), chi, edof, nn)
    fixed = plate.copy()
    fixed[arc] = True
    g = np.zeros(nn)
    g[arc] = b
    u = fem_direct.lifted_solve(K, np.zeros(nn), g, fixed)
    t = u if cfg.time_map == "identity" else heat_fields.poisson_distance(
        u, np.ones(nelx * nely), plate, nelx, nely)
    te = t[edof].mean(-1)
    part = x >= x.max() / 2
    return (te / te[part].max()).reshape(nely, nelx), u, t


@pytest.mark.parametrize("time_map, notch, e", [("identity", False, 0.0),
                                                ("identity", True, 3.0),
                                                ("poisson", True, 0.0)])
def test_laplace_time_field_matches_oracle(dev, time_map, notch, e):
    from sttopt_amd import virtual_heat

    nelx, nely = 36, 22
    active = np.ones((nely, nelx), bool)
    if notch:
        active[:8, 24:] = False
    mesh = virtual_heat.ScalarMesh.build(nelx, nely, "bottom_edge", dev, active=active)
    rng = np.random.default_rng(4)
    xPhys = rng.uniform(0.1, 1, (nely, nelx)) * active
    mu = rng.uniform(0.3, 0.7, (nely, nelx))
    wall = virtual_heat.distance_wall(mesh)
    H = virtual_heat.wall_filter(len(mesh.arc), 3.0)
    b = _t(H, dev) @ (wall + 0.05 * _t(rng.standard_normal(len(mesh.arc)), dev))
    cfg = types.SimpleNamespace(chi_contrast=10.0, chi_density_floor=1e-2,
                                chi_density_sensitivity=True, time_map=time_map,
                                laplace_cg_rtol=1e-12, poisson_cg_rtol=1e-12)
    sol = virtual_heat.laplace_time_field(cfg, mesh, _t(xPhys, dev), _t(mu, dev), b,
                                          density_exponent=e)
    want, u, t = _oracle_laplace(nelx, nely, active, mesh.plate.cpu().numpy(),
                                 mesh.arc.cpu().numpy(), xPhys, mu, b.cpu().numpy(), cfg, e)
    assert (sol.poisson is not None) == (time_map == "poisson")
    np.testing.assert_allclose(sol.primary.detach().cpu().numpy(), u, atol=1e-9)
    np.testing.assert_allclose(sol.t.detach().cpu().numpy(), t, atol=1e-8 * np.abs(t).max())
    np.testing.assert_allclose(sol.tPhys.detach().cpu().numpy(), want, atol=1e-8)
    # warm start from the previous solution: the same field
    sol2 = virtual_heat.laplace_time_field(cfg, mesh, _t(xPhys, dev), _t(mu, dev), b,
                                           previous=sol, density_exponent=e)
    np.testing.assert_allclose(sol2.tPhys.detach().cpu().numpy(), want, atol=1e-8)


@pytest.mark.parametrize("time_map", ["identity", "poisson"])
def test_laplace_time_field_gradient_matches_finite_differences(dev, time_map):
    """Autograd through the lifted solve (a real adjoint solve: the load depends on the
    wall data through K g) and, for the Poisson map, the direction field and the second
    solve, against central differences of the oracle along seeded directions in mu,
    xPhys (density exponent 2) and the wall data b."""
    from sttopt_amd import virtual_heat

    nelx, nely = 18, 12
    active = np.ones((nely, nelx), bool)
    active[:4, 12:] = False
    mesh = virtual_heat.ScalarMesh.build(nelx, nely, "bottom_edge", dev, active=active)
    plate, arc = mesh.plate.cpu().numpy(), mesh.arc.cpu().numpy()
    rng = np.random.default_rng(8)
    xPhys = rng.uniform(0.2, 1, (nely, nelx))
    mu = rng.uniform(0.3, 0.7, (nely, nelx))
    b = 0.5 * virtual_heat.distance_wall(mesh).cpu().numpy() + 0.02 * rng.standard_normal(arc.size)
    c = rng.standard_normal((nely, nelx))
    cfg = types.SimpleNamespace(chi_contrast=10.0, chi_density_floor=1e-2,
                                chi_density_sensitivity=True, time_map=time_map,
                                laplace_cg_rtol=1e-13, poisson_cg_rtol=1e-13)
    e = 2.0
    xt = _t(xPhys, dev).requires_grad_(True)
    mt = _t(mu, dev).requires_grad_(True)
    bt = _t(b, dev).requires_grad_(True)
    sol = virtual_heat.laplace_time_field(cfg, mesh, xt, mt, bt, density_exponent=e)
    (sol.tPhys * _t(c, dev)).sum().backward()
    edof = fem_direct.element_dofs(nelx, nely, 1)
    _, _, t0 = _oracle_laplace(nelx, nely, active, plate, arc, xPhys, mu, b, cfg, e)
    te0 = t0[edof].mean(-1)
    s0 = te0[xPhys.ravel() >= xPhys.max() / 2].max()  # the normalization is held constant

    def L(xv, mv, bv):
        _, _, t = _oracle_laplace(nelx, nely, active, plate, arc, xv, mv, bv, cfg, e)
        return ((t[edof].mean(-1).reshape(nely, nelx) / s0) * c).sum()

    h = 1e-6
    for k in range(3):
        dx = rng.standard_normal(xPhys.shape) * 0.1
        dm = rng.standard_normal(mu.shape)
        db = rng.standard_normal(b.shape) * 0.1
        fd = (L(xPhys + h * dx, mu + h * dm, b + h * db)
              - L(xPhys - h * dx, mu - h * dm, b - h * db)) / (2 * h)
        ad = float((xt.grad.cpu().numpy() * dx).sum() + (mt.grad.cpu().numpy() * dm).sum()
                   + (bt.grad.cpu().numpy() * db).sum())
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-8), (k, ad, fd)

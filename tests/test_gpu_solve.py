"""The differentiable solve (FemSolve) and the fused space-time compliance on the
device, against the assembled fp64 oracle, hand-derived sensitivities and central
finite differences (the protocol of the reference's tests/test_torch_solve.py and
tests/test_compliance.py:416-604)."""

import numpy as np
import pytest
import torch

from conftest import rel_err
from oracle import fem_direct

pytestmark = pytest.mark.gpu

EMIN, EMAX, PENAL, NU = 1e-9, 1.0, 3.0, 0.3


def _setup(nelx, nely, dev, case="cantilever"):
    from sttopt_amd import fem

    F, free = fem.load_case(case, nelx, nely)
    ndof = F.size
    mask = torch.zeros(ndof, dtype=torch.bool, device=dev)
    mask[torch.as_tensor(free, device=dev)] = True
    KE = torch.tensor(fem_direct.plane_stress_KE(NU), dtype=torch.float64, device=dev)
    edof = torch.tensor(fem_direct.element_dofs(nelx, nely), device=dev)
    return torch.tensor(F, device=dev), ndof, KE, edof, mask, free


def test_gradcheck_at_moderate_density(dev):
    from sttopt_amd import torch_solve

    nelx, nely = 4, 3
    F, ndof, KE, edof, mask, _ = _setup(nelx, nely, dev)
    rng = np.random.default_rng(0)
    density = torch.tensor(rng.uniform(0.2, 0.8, nelx * nely), dtype=torch.float64, device=dev,
                           requires_grad=True)
    F = F.clone().requires_grad_(True)

    def fn(d, f):
        return torch_solve.femsolve(d, f, edof, KE, mask, nelx, nely, rtol=1e-10)

    assert torch.autograd.gradcheck(fn, (density, F), eps=1e-6, atol=1e-5, rtol=1e-3)


@pytest.mark.parametrize("load_depends_on_density", [False, True])
def test_self_adjoint_backward_takes_zero_iterations(dev, load_depends_on_density):
    from sttopt_amd import compliance, torch_fem, torch_solve

    nelx, nely = 6, 4
    F0, ndof, KE, edof, mask, _ = _setup(nelx, nely, dev)
    rng = np.random.default_rng(2)
    density = torch.tensor(rng.uniform(0.1, 0.9, nelx * nely), dtype=torch.float64, device=dev,
                           requires_grad=True)
    if load_depends_on_density:
        w = torch.tensor(rng.standard_normal(ndof), dtype=torch.float64, device=dev)
        F = w * density.sum()
    else:
        F = F0
    info = {}
    U = torch_solve.femsolve(density, F, edof, KE, mask, nelx, nely, rtol=1e-12, info=info)
    ce = compliance._element_strain_energy(U, edof, KE, nely, nelx).reshape(-1)
    L = torch.sum(density * ce)
    g = torch.autograd.grad(L, U, retain_graph=True)[0]
    np.testing.assert_allclose(torch_fem.project(g, mask).detach().cpu().numpy(),
                               torch_fem.project(2 * F, mask).detach().cpu().numpy(),
                               rtol=1e-8, atol=1e-10)
    L.backward()
    assert info["backward_n_iter"] == 0


def test_backward_reuses_forward_hierarchy(dev):
    from sttopt_amd import torch_mg, torch_solve

    nelx, nely = 18, 12
    F, ndof, KE, edof, mask, _ = _setup(nelx, nely, dev)
    density = torch.tensor(np.random.default_rng(6).uniform(0.05, 1, nelx * nely), device=dev,
                           requires_grad=True)
    calls = 0
    orig = torch_mg.build_hierarchy

    def counting(*a, **k):
        nonlocal calls
        calls += 1
        return orig(*a, **k)

    torch_mg.build_hierarchy = counting
    try:
        U = torch_solve.femsolve(density, F, edof, KE, mask, nelx, nely, rtol=1e-10,
                                 max_coarse_elements=24)
        assert calls == 1
        torch.sum(U**2).backward()
        assert calls == 1
    finally:
        torch_mg.build_hierarchy = orig


def test_batched_matches_sequential_forward_and_backward(dev):
    from sttopt_amd import torch_solve

    nelx, nely, B = 18, 12, 5
    F, ndof, KE, edof, mask, _ = _setup(nelx, nely, dev)
    rng = np.random.default_rng(4)
    db = torch.tensor(rng.uniform(0.05, 1, (B, nelx * nely)), device=dev, requires_grad=True)
    Fb = F.expand(B, -1).clone().requires_grad_(True)
    kw = dict(rtol=1e-11, max_coarse_elements=24)
    Ub = torch_solve.femsolve(db, Fb, edof, KE, mask, nelx, nely, **kw)
    for s in range(B):
        ds = db[s].detach().clone().requires_grad_(True)
        fs = Fb[s].detach().clone().requires_grad_(True)
        Us = torch_solve.femsolve(ds, fs, edof, KE, mask, nelx, nely, **kw)
        np.testing.assert_allclose(Ub[s].detach().cpu().numpy(), Us.detach().cpu().numpy(),
                                   rtol=1e-6, atol=1e-6 * float(Us.detach().abs().max()))
        seed = torch.zeros_like(Ub)
        seed[s, 3] = 1.0
        gd, gf = torch.autograd.grad(Ub, (db, Fb), grad_outputs=seed, retain_graph=True)
        seed1 = torch.zeros_like(Us)
        seed1[3] = 1.0
        gds, gfs = torch.autograd.grad(Us, (ds, fs), grad_outputs=seed1)
        np.testing.assert_allclose(gd[s].cpu().numpy(), gds.cpu().numpy(), rtol=1e-6,
                                   atol=1e-9 * float(gds.abs().max()))
        np.testing.assert_allclose(gf[s].cpu().numpy(), gfs.cpu().numpy(), rtol=1e-6,
                                   atol=1e-9 * float(gfs.abs().max()))
        others = [i for i in range(B) if i != s]
        assert torch.all(gd[others] == 0) and torch.all(gf[others] == 0)


def test_nonconvergence_propagates_from_forward_and_backward(dev):
    from sttopt_amd import torch_fem, torch_solve

    nelx, nely = 12, 8
    F, ndof, KE, edof, mask, _ = _setup(nelx, nely, dev)
    kw = dict(max_coarse_elements=24)
    density = torch.full((nelx * nely,), 0.4, dtype=torch.float64, device=dev, requires_grad=True)
    with pytest.raises(torch_fem.CGConvergenceError):
        torch_solve.femsolve(density, F, edof, KE, mask, nelx, nely, rtol=1e-14, max_iter=2, **kw)
    # an exact warm start converges before the first iteration even at max_iter=0;
    # a non-compliance objective then needs a real adjoint solve, which must raise
    U = torch_solve.femsolve(density.detach(), F, edof, KE, mask, nelx, nely, rtol=1e-12, **kw)
    U0 = torch_solve.femsolve(density, F, edof, KE, mask, nelx, nely, rtol=1e-8, max_iter=0,
                              x0=U.detach(), **kw)
    w = torch.linspace(0, 1, ndof, dtype=torch.float64, device=dev)
    with pytest.raises(torch_fem.CGConvergenceError):
        (U0 * w).sum().backward()


def test_lifted_scalar_solve_matches_direct(dev):
    """Scalar heat solve with a drain and Dirichlet values (the config-3 primitive)."""
    from sttopt_amd import virtual_heat

    nelx, nely = 90, 30
    rng = np.random.default_rng(4)
    mesh = virtual_heat.ScalarMesh.build(nelx, nely, "left_edge", dev)
    chi = 10 ** rng.uniform(-3, 0, nelx * nely)
    react = np.full(nelx * nely, 0.05)
    nn = mesh.ndof
    cols = np.arange(nn) % (nelx + 1)
    fixed = cols == 0
    assert np.array_equal(mesh.plate.cpu().numpy(), fixed)
    F = rng.uniform(-1, 1, nn)
    g = np.where(fixed, rng.uniform(0, 1, nn), 0.0)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)  # noqa: E731
    u = mesh.solve(t(chi), t(g), torch.tensor(fixed, device=dev), 1e-11, F=t(F), reaction=t(react))
    K = fem_direct.assemble(fem_direct.diffusion_KE(), chi, fem_direct.element_dofs(nelx, nely, 1),
                            nn, reaction=react)
    want = fem_direct.lifted_solve(K, F, g, fixed)
    assert np.abs(u.cpu().numpy() - want).max() <= 1e-7 * np.abs(want).max()


def test_space_time_compliance_matches_oracle(dev):
    """Fused batched compliance: values and hand-written adjoint vs the assembled
    oracle with hand-derived sensitivities."""
    from sttopt_amd import compliance, workloads

    p = workloads.space_time(48, 30, 4, beta=16.0, seed=7, device=dev)
    x = p.xPhys.clone().requires_grad_(True)
    t = p.tPhys.clone().requires_grad_(True)
    c, stages, U = compliance.batched_whole_and_gravity_compliance(
        x, t, p.KE, p.edofMat, EMIN, EMAX, PENAL, p.freedofs, p.F, p.ndof, p.C, p.beta,
        p.stage_times)
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(p.nelx, p.nely)
    c_ref, dcx_ref, st_ref, U_ref = fem_direct.space_time(
        p.xPhys.cpu().numpy(), p.tPhys.cpu().numpy(), p.stage_times, p.beta, KE, edof,
        p.F.cpu().numpy(), p.freedofs.cpu().numpy(), p.C)
    assert abs(float(c.detach()) - c_ref) <= 1e-6 * abs(c_ref)
    for k, (cg, dx, dt) in enumerate(st_ref):
        assert abs(float(stages[k]) - cg) <= 1e-6 * abs(cg)
    assert rel_err(U.cpu().numpy(), U_ref) < 1e-6
    # per-row gradients through the hand-written adjoint
    rows = [c] + list(stages)
    refs = [(dcx_ref, np.zeros_like(dcx_ref))] + [(s[1], s[2]) for s in st_ref]
    for out, (gx_ref, gt_ref) in zip(rows, refs):
        gx, gt = torch.autograd.grad(out, (x, t), retain_graph=True)
        assert rel_err(gx.cpu().numpy().ravel(), gx_ref) < 1e-5
        if np.linalg.norm(gt_ref) > 0:
            assert rel_err(gt.cpu().numpy().ravel(), gt_ref) < 1e-5
        else:
            assert float(gt.abs().max()) == 0.0


@pytest.mark.parametrize("composed", [False, True])
def test_jacobian_rows_through_the_space_time_compliance(dev, composed):
    """`jacobian_rows` of the 1 + nStage compliances w.r.t. (x, t, an unused leaf), the
    way the drivers call it: each row equals the oracle's hand-derived gradients of that
    row alone (several backward passes through one retained graph), and the unused
    leaf's block is exactly zero.  composed: through the reference composition (FemSolve
    autograd, selected by patching `_solve_fe_batched`) instead of the fused Function."""
    from sttopt_amd import compliance, sensitivity, workloads

    p = workloads.space_time(36, 20, 3, beta=16.0, seed=4, device=dev)
    x = p.xPhys.clone().requires_grad_(True)
    t = p.tPhys.clone().requires_grad_(True)
    spare = torch.ones(5, dtype=torch.float64, device=dev, requires_grad=True)
    orig = compliance._solve_fe_batched
    C = p.C  # the fused path takes the stencil weight; the composition needs the matrix
    if composed:
        compliance._solve_fe_batched = lambda *a, **k: orig(*a, **k)
        Cs = fem_direct.gravity_matrix(p.nelx, p.nely).tocoo()
        C = torch.sparse_coo_tensor(np.vstack([Cs.row, Cs.col]), Cs.data, Cs.shape,
                                    device=dev).coalesce()
    try:
        c, stages, _ = compliance.batched_whole_and_gravity_compliance(
            x, t, p.KE, p.edofMat, EMIN, EMAX, PENAL, p.freedofs, p.F, p.ndof, C, p.beta,
            p.stage_times)
        jx, jt, js = sensitivity.jacobian_rows(torch.stack([c] + list(stages)), (x, t, spare))
    finally:
        compliance._solve_fe_batched = orig
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(p.nelx, p.nely)
    _, dcx_ref, st_ref, _ = fem_direct.space_time(
        p.xPhys.cpu().numpy(), p.tPhys.cpu().numpy(), p.stage_times, p.beta, KE, edof,
        p.F.cpu().numpy(), p.freedofs.cpu().numpy(), p.C)
    k = 1 + len(p.stage_times)
    assert jx.shape == (k, p.nel) and jt.shape == (k, p.nel) and js.shape == (k, 5)
    assert float(js.abs().max()) == 0.0
    assert rel_err(jx[0].cpu().numpy(), dcx_ref) < 1e-5
    assert float(jt[0].abs().max()) == 0.0  # the whole structure does not see t
    for i, (_, gx_ref, gt_ref) in enumerate(st_ref, start=1):
        assert rel_err(jx[i].cpu().numpy(), gx_ref) < 1e-5, i
        assert rel_err(jt[i].cpu().numpy(), gt_ref) < 1e-5, i


def test_space_time_fused_equals_composed_autograd(dev):
    """The fused path against the reference composition (torch algebra + FemSolve
    autograd), selected by patching the `_solve_fe_batched` hook."""
    from sttopt_amd import compliance, workloads

    p = workloads.space_time(40, 24, 3, beta=16.0, seed=2, device=dev)
    args = (p.KE, p.edofMat, EMIN, EMAX, PENAL, p.freedofs, p.F, p.ndof)
    x1 = p.xPhys.clone().requires_grad_(True)
    t1 = p.tPhys.clone().requires_grad_(True)
    c1, s1, _ = compliance.batched_whole_and_gravity_compliance(x1, t1, *args, p.C, p.beta, p.stage_times)
    (c1 + 2 * sum(s1)).backward()
    import scipy.sparse as sp

    C = fem_direct.gravity_matrix(p.nelx, p.nely).tocoo()
    Ct = torch.sparse_coo_tensor(np.vstack([C.row, C.col]), C.data, C.shape, device=dev).coalesce()
    orig = compliance._solve_fe_batched
    compliance._solve_fe_batched = lambda *a, **k: orig(*a, **k)
    try:
        x2 = p.xPhys.clone().requires_grad_(True)
        t2 = p.tPhys.clone().requires_grad_(True)
        c2, s2, _ = compliance.batched_whole_and_gravity_compliance(x2, t2, *args, Ct, p.beta,
                                                                     p.stage_times)
        (c2 + 2 * sum(s2)).backward()
    finally:
        compliance._solve_fe_batched = orig
    assert abs(float(c1 - c2)) <= 1e-7 * abs(float(c2))
    assert rel_err(x1.grad.cpu().numpy(), x2.grad.cpu().numpy()) < 1e-6
    assert rel_err(t1.grad.cpu().numpy(), t2.grad.cpu().numpy()) < 1e-6
    assert sp is not None


def test_space_time_gradient_matches_central_differences(dev):
    """Hand-written adjoint vs central FD (h = 1e-6) through the whole fused solve on
    seeded elements (the C2 protocol at a reduced mesh, print-order time field).  The
    differenced compliances are solved to rtol 1e-11 with the fp64 V-cycle: at 1e-8
    the solve error, not the adjoint, would dominate the difference."""
    from sttopt_amd import compliance, torch_mg, torch_solve, workloads

    old = (torch_solve.DEFAULT_CG_RTOL, torch_mg.VCYCLE_DTYPE)
    torch_solve.DEFAULT_CG_RTOL, torch_mg.VCYCLE_DTYPE = 1e-11, torch.float64
    try:
        _fd_check(compliance, workloads, dev)
    finally:
        torch_solve.DEFAULT_CG_RTOL, torch_mg.VCYCLE_DTYPE = old


def _fd_check(compliance, workloads, dev):

    p = workloads.space_time(60, 30, 10, beta=16.0, seed=1, time_field="print", device=dev)
    args = (p.KE, p.edofMat, EMIN, EMAX, PENAL, p.freedofs, p.F, p.ndof, p.C, p.beta, p.stage_times)

    def total(x, t):
        c, st, _ = compliance.batched_whole_and_gravity_compliance(x, t, *args)
        return c + sum(st)

    x = p.xPhys.clone().requires_grad_(True)
    t = p.tPhys.clone().requires_grad_(True)
    total(x, t).backward()
    rng = np.random.default_rng(11)
    elems = rng.choice(p.nel, 20, replace=False)
    h = 1e-6
    worst_x = worst_t = 0.0
    with torch.no_grad():
        for e in elems:
            for which, grad in (("x", x.grad), ("t", t.grad)):
                vals = []
                for sgn in (1, -1):
                    xx, tt = p.xPhys.clone(), p.tPhys.clone()
                    (xx if which == "x" else tt).view(-1)[e] += sgn * h
                    vals.append(float(total(xx, tt)))
                fd = (vals[0] - vals[1]) / (2 * h)
                g = float(grad.view(-1)[e])
                scale = float(grad.abs().max())
                err = abs(fd - g) / max(abs(g), 1e-3 * scale)
                if which == "x":
                    worst_x = max(worst_x, err)
                else:
                    worst_t = max(worst_t, err)
    # the t-derivatives of far-from-front elements are ~exp(-2 beta |t - t_k|) small:
    # their central differences carry more cancellation noise
    assert worst_x < 1e-4 and worst_t < 1e-3, (worst_x, worst_t)

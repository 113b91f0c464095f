"""Operator-level parity of the HIP kernels against the assembled fp64 oracle
(the protocol of the reference's tests/test_torch_fem.py:54-141): matvec, diagonal,
projected operator, lumped reaction, coarse (keff) apply, grid transfers."""

import numpy as np
import pytest
import torch

from conftest import density_fields
from oracle import fem_direct, mgcg_port

pytestmark = pytest.mark.gpu

EMIN, EMAX, PENAL, NU = 1e-9, 1.0, 3.0, 0.3
MESHES = [(4, 3), (12, 8), (30, 10), (120, 40)]


def _t(a, dev, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=dev)


def _cantilever(nelx, nely):
    from sttopt_amd import fem

    return fem.load_case("cantilever", nelx, nely)


@pytest.mark.parametrize("nelx, nely", MESHES)
def test_matvec_and_diagonal_match_assembled_K(dev, nelx, nely):
    from sttopt_amd import torch_fem

    rng = np.random.default_rng(0)
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    ndof = 2 * (nelx + 1) * (nely + 1)
    for name, x in density_fields(nelx, nely, rng).items():
        dens = fem_direct.simp(x, EMIN, EMAX, PENAL)
        K = fem_direct.assemble(KE, dens, edof, ndof)
        v = rng.standard_normal(ndof)
        got = torch_fem.matvec(_t(v, dev), _t(dens, dev), _t(edof, dev, torch.int64), _t(KE, dev), ndof)
        want = K @ v
        np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
        diag = torch_fem.matvec_diagonal(_t(dens, dev), _t(edof, dev, torch.int64), _t(KE, dev), ndof)
        np.testing.assert_allclose(diag.cpu().numpy(), K.diagonal(), rtol=1e-10, atol=0)


@pytest.mark.parametrize("nelx, nely", [(12, 8), (30, 10)])
def test_operator_projection_and_batching(dev, nelx, nely):
    from sttopt_amd import torch_fem

    rng = np.random.default_rng(1)
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    ndof = 2 * (nelx + 1) * (nely + 1)
    _, free = _cantilever(nelx, nely)
    mask = np.zeros(ndof, bool)
    mask[free] = True
    B = 5
    dens = rng.uniform(0.05, 1.0, (B, nelx * nely))
    v = rng.standard_normal((B, ndof))
    rdiag = rng.uniform(0, 0.3, (B, ndof))
    got = torch_fem.operator(_t(v, dev), _t(dens, dev), _t(edof, dev, torch.int64), _t(KE, dev), ndof,
                             torch.as_tensor(mask, device=dev), _t(rdiag, dev)).cpu().numpy()
    for b in range(B):
        K = fem_direct.assemble(KE, dens[b], edof, ndof)
        want = (K @ (v[b] * mask) + rdiag[b] * v[b] * mask) * mask
        np.testing.assert_allclose(got[b], want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
        assert np.all(got[b][~mask] == 0.0)
    # shared density, batched vectors (stride-0 rows)
    got = torch_fem.matvec(_t(v, dev), _t(dens[0], dev), _t(edof, dev, torch.int64), _t(KE, dev), ndof)
    K = fem_direct.assemble(KE, dens[0], edof, ndof)
    np.testing.assert_allclose(got.cpu().numpy(), (K @ v.T).T, rtol=1e-10, atol=1e-12)
    # diag with mask: 1 at fixed dofs
    dg = torch_fem.jacobi_preconditioner_diag(_t(dens[0], dev), _t(edof, dev, torch.int64), _t(KE, dev),
                                              ndof, torch.as_tensor(mask, device=dev)).cpu().numpy()
    np.testing.assert_allclose(dg[mask], K.diagonal()[mask], rtol=1e-12)
    assert np.all(dg[~mask] == 1.0)


def test_fp32_matvec_close_to_fp64(dev):
    from sttopt_amd import torch_fem

    nelx, nely = 64, 48
    rng = np.random.default_rng(2)
    KE = _t(fem_direct.plane_stress_KE(NU), dev)
    edof = _t(fem_direct.element_dofs(nelx, nely), dev, torch.int64)
    ndof = 2 * (nelx + 1) * (nely + 1)
    dens = _t(rng.uniform(0.1, 1, nelx * nely), dev)
    v = _t(rng.standard_normal((3, ndof)), dev)
    y64 = torch_fem.matvec(v, dens, edof, KE, ndof)
    y32 = torch_fem.matvec(v.float(), dens.float(), edof, KE, ndof)
    assert y32.dtype == torch.float32
    err = (y32.double() - y64).norm() / y64.norm()
    assert err < 1e-6, err


@pytest.mark.parametrize("nelx, nely", [(6, 4), (30, 10)])
def test_scalar_operator_with_reaction(dev, nelx, nely):
    from sttopt_amd import torch_fem

    rng = np.random.default_rng(3)
    KE = fem_direct.diffusion_KE()
    edof = fem_direct.element_dofs(nelx, nely, 1)
    ndof = (nelx + 1) * (nely + 1)
    chi = rng.uniform(0.01, 1, nelx * nely)
    react = rng.uniform(0, 0.5, nelx * nely)
    K = fem_direct.assemble(KE, chi, edof, ndof, reaction=react)
    rd = torch_fem.lumped_reaction(_t(react, dev), _t(edof, dev, torch.int64), ndof)
    diag_r = np.zeros(ndof)
    np.add.at(diag_r, edof.ravel(), np.repeat(react / 4, 4))
    np.testing.assert_allclose(rd.cpu().numpy(), diag_r, rtol=1e-14, atol=1e-15)
    mask = np.ones(ndof, bool)
    v = rng.standard_normal(ndof)
    got = torch_fem.operator(_t(v, dev), _t(chi, dev), _t(edof, dev, torch.int64), _t(KE, dev), ndof,
                             torch.as_tensor(mask, device=dev), rd)
    np.testing.assert_allclose(got.cpu().numpy(), K @ v, rtol=1e-10, atol=1e-12)


def test_edof_validation_rejects_non_grid_maps(dev):
    from sttopt_amd import torch_fem

    nelx, nely = 6, 4
    edof = fem_direct.element_dofs(nelx, nely)
    bad = edof.copy()
    bad[3, [0, 1]] = bad[3, [1, 0]]
    ndof = 2 * (nelx + 1) * (nely + 1)
    with pytest.raises(ValueError):
        torch_fem.matvec(torch.zeros(ndof, device=dev, dtype=torch.float64),
                         torch.ones(nelx * nely, device=dev, dtype=torch.float64),
                         torch.as_tensor(bad, device=dev), _t(fem_direct.plane_stress_KE(NU), dev), ndof)


@pytest.mark.parametrize("kx, ky", [(1, 1), (2, 2), (3, 3), (3, 1), (2, 3)])
def test_hip_transfers_match_node_grid_transfers(dev, kx, ky):
    """The HIP restriction / prolongation equal the torch node-grid transfers and
    are exact adjoints (<P c, f> == <c, R f>)."""
    from sttopt_amd import _lib, torch_mg

    ncx, ncy = 5, 3
    rng = np.random.default_rng(4)
    for d in (1, 2):
        c = _t(rng.standard_normal(d * (ncx + 1) * (ncy + 1)), dev)
        f = _t(rng.standard_normal(d * (ncx * kx + 1) * (ncy * ky + 1)), dev)
        Pc_ref = mgcg_port.on_node_grid(c.cpu(), ncx + 1, ncy + 1, kx, ky, mgcg_port.interp_axis)
        Rf_ref = mgcg_port.on_node_grid(f.cpu(), ncx * kx + 1, ncy * ky + 1, kx, ky, mgcg_port.restrict_axis)
        Rf = torch.empty_like(c)
        lib = _lib.load()
        s = _lib.stream_handle(dev)
        _lib.check(lib.sttopt_restrict(_lib.F64, d, ncx, ncy, kx, ky, 1, f.data_ptr(), 0, None,
                                       Rf.data_ptr(), 0, s), "restrict")
        Pc = torch.zeros_like(f)
        _lib.check(lib.sttopt_prolong_add(_lib.F64, d, ncx, ncy, kx, ky, 1, c.data_ptr(), 0, None,
                                          Pc.data_ptr(), 0, s), "prolong")
        np.testing.assert_allclose(Rf.cpu().numpy(), Rf_ref.numpy(), rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(Pc.cpu().numpy(), Pc_ref.numpy(), rtol=1e-13, atol=1e-13)
        lhs = float((Pc * f).sum())
        rhs = float((c * Rf).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
        # the package's torch helpers agree as well
        Pc2 = torch_mg._on_node_grid(c, ncx + 1, ncy + 1, kx, ky, torch_mg._interp_axis)
        Rf2 = torch_mg._on_node_grid(f, ncx * kx + 1, ncy * ky + 1, kx, ky, torch_mg._restrict_axis)
        np.testing.assert_allclose(Pc2.cpu().numpy(), Pc_ref.numpy(), rtol=1e-14, atol=1e-14)
        np.testing.assert_allclose(Rf2.cpu().numpy(), Rf_ref.numpy(), rtol=1e-14, atol=1e-14)


def test_device_required():
    from sttopt_amd import _lib, torch_fem

    with pytest.raises(_lib.DeviceError):
        torch_fem.matvec(torch.zeros(30), torch.ones(6), torch.zeros(6, 8, dtype=torch.int64),
                         torch.eye(8), 30)


def test_projected_operator_mode_equals_operator_on_projected_inputs(dev):
    """STTOPT_MODE_OPERATOR_PROJECTED (the outer PCG's apply: no input projection) gives
    the operator's result bit for bit on inputs that are zero at the fixed dofs, with
    and without a reaction diagonal, and the assembled P K P v."""
    from sttopt_amd import _grid, _lib

    nelx, nely, B = 70, 20, 3
    rng = np.random.default_rng(6)
    KE = fem_direct.plane_stress_KE(NU)
    edof = fem_direct.element_dofs(nelx, nely)
    ndof = 2 * (nelx + 1) * (nely + 1)
    _, free = _cantilever(nelx, nely)
    mask = np.zeros(ndof, bool)
    mask[free] = True
    dens = rng.uniform(0.05, 1.0, (B, nelx * nely))
    v = rng.standard_normal((B, ndof)) * mask
    bits = _grid.mask_bits(torch.as_tensor(mask, device=dev))
    lib = _lib.load()
    stream = torch.cuda.current_stream(dev).cuda_stream
    for rd in (None, rng.uniform(0, 0.3, (B, ndof))):
        outs = []
        for mode in (_lib.MODE_OPERATOR, _lib.MODE_OPERATOR_PROJECTED):
            vt, dt = _t(v, dev), _t(dens, dev)
            rt = None if rd is None else _t(rd, dev)
            y = torch.empty_like(vt)
            _lib.check(lib.sttopt_apply(_lib.F64, mode, 2, nelx, nely, B, vt.data_ptr(), ndof,
                                        dt.data_ptr(), nelx * nely, _lib.ke_array(_t(KE, dev)),
                                        bits.data_ptr(), None if rt is None else rt.data_ptr(),
                                        0 if rt is None else ndof, y.data_ptr(), ndof, stream),
                       "sttopt_apply")
            outs.append(y.cpu().numpy())
        assert np.array_equal(outs[0], outs[1])
        K = fem_direct.assemble(KE, dens[0], edof, ndof)
        want = (K @ v[0] + (0 if rd is None else rd[0] * v[0])) * mask
        np.testing.assert_allclose(outs[1][0], want, rtol=1e-10, atol=1e-12 * np.abs(want).max())

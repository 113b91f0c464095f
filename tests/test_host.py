"""CPU-side checks: the C-ABI library loads and exports every declared entry point;
host logic (coarsening, transfers, interpolation matrices, workloads, load cases);
every kernel-backed entry point refuses CPU tensors (there is no CPU path)."""

import ctypes
import re
import subprocess
import sys

import numpy as np
import pytest
import torch

from conftest import ROOT
from oracle import mgcg_port


def _declared_functions():
    text = (ROOT / "include" / "sttopt_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sttopt_\w+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    from sttopt_amd import _lib

    lib = _lib.load()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == _lib.exported_symbols()
    assert lib.sttopt_version() >= 1
    lb, mb = ctypes.c_int64(), ctypes.c_int64()
    assert lib.sttopt_abi_sizes(ctypes.byref(lb), ctypes.byref(mb)) == 0
    assert lb.value == ctypes.sizeof(_lib.Level) and mb.value == ctypes.sizeof(_lib.MG)


def test_library_reports_errors_without_touching_the_gpu():
    from sttopt_amd import _lib

    lib = _lib.load()
    st = lib.sttopt_apply(_lib.F64, 0, 3, 4, 4, 1, None, 0, None, 0, None, None, None, 0, None, 0, None)
    assert st == _lib.EINVAL and "invalid grid" in _lib.last_error()
    st = lib.sttopt_galerkin(2, 5, 4, 2, 2, 1, None, 0, None, 0, None, None, None, 0, None, None, 0, None)
    assert st == _lib.EINVAL
    # d = 2 nodes are moved as 2-vectors: a pointer off the 16-byte node pair, or an odd
    # row stride, is refused before any launch (the addresses are never dereferenced)
    ke = _lib.ke_array(torch.eye(8, dtype=torch.float64))
    ndof = 2 * 5 * 5
    st = lib.sttopt_apply(_lib.F64, 0, 2, 4, 4, 1, 4096 + 8, ndof, 8192, 16, ke, None, None, 0,
                          12288, ndof, None)
    assert st == _lib.EINVAL and "node-aligned" in _lib.last_error()
    st = lib.sttopt_apply(_lib.F64, 0, 2, 4, 4, 2, 4096, ndof + 1, 8192, 16, ke, None, None, 0,
                          12288, ndof, None)
    assert st == _lib.EINVAL and "node-aligned" in _lib.last_error()
    # the element gathers of the gradient kernels load nodes as 2-vectors too
    st = lib.sttopt_strain_energy(2, 4, 4, 1, 4096 + 8, ndof, ke, 8192, 16, None)
    assert st == _lib.EINVAL and "node-aligned" in _lib.last_error()
    st = lib.sttopt_adjoint_grad(2, 4, 4, 2, 4096, ndof, 8192, ndof + 1, ke, 12288, 16, None, 0,
                                 None)
    assert st == _lib.EINVAL and "node-aligned" in _lib.last_error()
    st = lib.sttopt_stage_epilogue(4, 4, 1, 4096, 8192, 12288, 16.0, 1e-9, 1.0, 3.0, 1.0, ke,
                                   16384 + 8, 20480, None, 24576, None, None, 28672, None)
    assert st == _lib.EINVAL and "node-aligned" in _lib.last_error()
    # the table-derived coarse diagonal takes at most 4 positions and d = 1 or 2
    table = (ctypes.c_double * 64)()
    st = lib.sttopt_param_diagonal(2, 4, 4, 3, 2, 1, 4096, 16, table, 8192, 12288, 50, None)
    assert st == _lib.EINVAL and "at most 4 positions" in _lib.last_error()
    st = lib.sttopt_param_diagonal(3, 4, 4, 2, 2, 1, 4096, 16, table, 8192, 12288, 50, None)
    assert st == _lib.EINVAL
    # the batched MG-PCG needs at least one group
    n_iter = ctypes.c_int32(0)
    st = lib.sttopt_mgpcg_batch(4096, 0, 1e-8, 10, 1, 1, ctypes.byref(n_iter), None)
    assert st == _lib.EINVAL and "no groups" in _lib.last_error()


def test_coarsening_factors():
    from sttopt_amd import torch_mg

    cases = {(180, 60): (2, 2), (45, 15): (3, 3), (15, 5): (3, 1), (5, 5): (1, 1), (1024, 512): (2, 2)}
    for (nx, ny), want in cases.items():
        assert torch_mg.coarsening_factors(nx, ny) == want
        assert mgcg_port.coarsening_factors(nx, ny) == want


@pytest.mark.parametrize("kx, ky", [(1, 1), (2, 2), (3, 3), (3, 1), (2, 3)])
def test_node_grid_transfers_are_adjoint_and_match_the_port(kx, ky):
    from sttopt_amd import torch_mg

    rng = np.random.default_rng(0)
    ncx, ncy = 3, 2
    c = torch.tensor(rng.standard_normal(2 * (ncx + 1) * (ncy + 1)))
    f = torch.tensor(rng.standard_normal(2 * (ncx * kx + 1) * (ncy * ky + 1)))
    Pc = torch_mg._on_node_grid(c, ncx + 1, ncy + 1, kx, ky, torch_mg._interp_axis)
    Rf = torch_mg._on_node_grid(f, ncx * kx + 1, ncy * ky + 1, kx, ky, torch_mg._restrict_axis)
    assert abs(float((Pc * f).sum() - (c * Rf).sum())) < 1e-12
    np.testing.assert_allclose(
        Pc.numpy(), mgcg_port.on_node_grid(c, ncx + 1, ncy + 1, kx, ky, mgcg_port.interp_axis).numpy(),
        atol=1e-15)
    np.testing.assert_allclose(
        Rf.numpy(),
        mgcg_port.on_node_grid(f, ncx * kx + 1, ncy * ky + 1, kx, ky, mgcg_port.restrict_axis).numpy(),
        atol=1e-14)


@pytest.mark.parametrize("kx, ky, d", [(2, 2, 2), (3, 3, 2), (2, 3, 1), (3, 1, 2)])
def test_child_interpolation_matrices(kx, ky, d):
    from sttopt_amd import torch_mg

    S = torch_mg.child_interpolation_matrices(kx, ky, None, torch.float64, d)
    np.testing.assert_allclose(S.numpy(), mgcg_port.child_matrices(kx, ky, d).numpy(), atol=1e-15)
    # interpolation reproduces constants: every row sums to 1 per component
    np.testing.assert_allclose(S.sum(-1).numpy(), np.ones(S.shape[:2]), atol=1e-15)


@pytest.mark.parametrize("nelx, nely, d", [(4, 3, 2), (5, 2, 1), (1, 1, 2)])
def test_element_dof_map_tensor_equals_the_explicit_map(nelx, nely, d):
    from oracle import fem_direct
    from sttopt_amd import fem, torch_mg

    t = torch_mg.element_dof_map_tensor(nelx, nely, "cpu", d)
    assert t.dtype == torch.int64
    np.testing.assert_array_equal(t.numpy(), fem_direct.element_dofs(nelx, nely, d))
    np.testing.assert_array_equal(t.numpy(), fem.element_dof_map(nelx, nely, d))


def test_child_elements_order():
    from sttopt_amd import torch_mg

    ch = torch_mg._child_elements(4, 2, 2, 2, "cpu")
    np.testing.assert_array_equal(ch.numpy(), [[0, 1, 4, 5], [2, 3, 6, 7]])
    np.testing.assert_array_equal(ch.numpy(), mgcg_port.child_elements(4, 2, 2, 2).numpy())


def test_workload_sizes_and_determinism():
    from sttopt_amd import workloads

    assert workloads.mbb(device="cpu").ndof == 9922
    a = workloads.space_time(240, 120, 10, device="cpu")
    b = workloads.space_time(240, 120, 10, device="cpu")
    assert a.ndof == 58322 and torch.equal(a.xPhys, b.xPhys) and torch.equal(a.tPhys, b.tPhys)
    assert float(a.tPhys.min()) == 0.0 and float(a.tPhys.max()) == 1.0
    assert a.stage_times[-1] == 1.0 and len(a.stage_times) == 10
    assert float(a.xPhys.min()) >= 1e-3 and float(a.xPhys.max()) <= 1.0
    h = workloads.heat(device="cpu")
    assert h.F.numel() == 115921 and abs(float(h.F.sum()) - 480 * 240) < 1e-6


def test_free_mask_project_and_safe_div_are_plain_tensor_algebra():
    from sttopt_amd import torch_fem

    m = torch_fem.free_mask(6, torch.tensor([0, 2, 5]))
    assert m.tolist() == [True, False, True, False, False, True]
    assert torch.equal(torch_fem.project(torch.ones(6), m), m.double())
    out = torch_fem.safe_div(torch.tensor([1.0, 2.0]), torch.tensor([0.0, 4.0]))
    assert out.tolist() == [0.0, 0.5]


def test_kernel_entry_points_refuse_cpu_tensors():
    from sttopt_amd import _lib, compliance, torch_fem, torch_mg, torch_solve, workloads

    p = workloads.space_time(8, 6, 2, device="cpu")
    dens = torch.ones(p.nel, dtype=torch.float64)
    with pytest.raises(_lib.DeviceError):
        torch_fem.matvec(torch.zeros(p.ndof, dtype=torch.float64), dens, p.edofMat, p.KE, p.ndof)
    with pytest.raises(_lib.DeviceError):
        torch_mg.build_hierarchy(dens, p.edofMat, p.KE, p.mask, p.nelx, p.nely)
    with pytest.raises(_lib.DeviceError):
        torch_solve.femsolve(dens, p.F, p.edofMat, p.KE, p.mask, p.nelx, p.nely)
    with pytest.raises(_lib.DeviceError):
        compliance.batched_whole_and_gravity_compliance(
            p.xPhys, p.tPhys, p.KE, p.edofMat, 1e-9, 1.0, 3.0, p.freedofs, p.F, p.ndof, p.C,
            p.beta, p.stage_times)
    with pytest.raises(_lib.DeviceError):
        torch_fem.pcg(lambda v: v, torch.ones(3), lambda r: r)


def test_bench_refuses_more_than_one_gpu():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2 and "one GPU" in out.stderr


def test_plate_nodes_and_wall_arc_on_a_rectangle_and_an_l_shape():
    """Host-side mesh data of the scalar mesh (virtual_heat.py:79-172): the plate is the
    print-base edge; the wall arc runs from the plate's end around the domain back to
    its start, clockwise, each off-plate boundary node once."""
    import numpy as np

    from sttopt_amd import virtual_heat as vh

    nelx, nely = 4, 3
    nn = (nelx + 1) * (nely + 1)
    node = lambda r, c: r * (nelx + 1) + c  # noqa: E731
    plate = vh.plate_nodes(nelx, nely, "bottom_edge")
    assert plate.sum() == nelx + 1 and all(plate[node(nely, c)] for c in range(nelx + 1))
    arc = vh.wall_arc(nelx, nely, "bottom_edge")
    want = ([node(r, 0) for r in range(nely - 1, -1, -1)] + [node(0, c) for c in range(1, nelx + 1)]
            + [node(r, nelx) for r in range(1, nely)])
    assert arc.tolist() == want
    assert vh.plate_nodes(nelx, nely, "left_edge").reshape(nely + 1, nelx + 1)[:, 0].all()
    with pytest.raises(ValueError):
        vh.plate_nodes(nelx, nely, "top_edge")
    # L-shape: the top-right 2 x 2 elements removed
    active = np.ones((nely, nelx), bool)
    active[:2, 2:] = False
    dom = vh.domain_nodes(active)
    assert not dom[0, 3] and dom[0, 2] and dom[2, 4]
    loop = vh._boundary_loop(active)
    assert len(set(loop.tolist())) == len(loop) == 14
    arc = vh.wall_arc(nelx, nely, "left_edge", active)
    lp = vh.plate_nodes(nelx, nely, "left_edge", active)
    assert not lp[arc].any() and len(arc) == len(loop) - lp.sum()
    assert arc[0] == node(0, 1) and arc[-1] == node(nely, 1)
    # a domain with two pieces is rejected
    two = np.zeros((nely, nelx), bool)
    two[0, 0] = two[2, 3] = True
    with pytest.raises(ValueError):
        vh._boundary_loop(two)
    assert vh.ScalarMesh.build(nelx, nely, "bottom_edge", "cpu").arc.numel() == nn - (nelx + 1) - (nelx - 1) * (nely - 1)


def test_density_filter_matrix_equals_oracle_and_continuity_filter_rows():
    """Host-side filter matrices (filters.py:38-106) against the oracle's pairwise
    construction; the continuity operator has unit diagonal and zero row sums."""
    import numpy as np
    import scipy.sparse as sp

    from oracle import filters_ref
    from sttopt_amd import filters, torch_util

    for nelx, nely, rmin in ((7, 5, 1.5), (9, 4, 2.3), (6, 6, 3.0), (5, 3, 1.0)):
        H, Hs = filters.density_filter(nelx, nely, rmin)
        Ho, Hso = filters_ref.filter_matrix(nelx, nely, rmin)
        np.testing.assert_allclose(H.toarray(), Ho, atol=1e-15)
        np.testing.assert_allclose(Hs, Hso, atol=1e-14)
        assert H._sttopt_stencil == (nelx, nely, float(rmin))
    L = filters.continuity_filter(8, 5, 2.0)
    np.testing.assert_allclose(L.diagonal(), 1.0)
    np.testing.assert_allclose(np.asarray(L.sum(axis=1)).ravel(), 0.0, atol=1e-14)
    with pytest.raises(ValueError):
        filters.continuity_filter(4, 3, 1.0)
    with pytest.raises(ValueError):
        torch_util.symmetric_csr_to_tensor(sp.csr_matrix(np.triu(np.ones((3, 3)))), "cpu", torch.float64)
    H, _ = filters.density_filter(5, 4, 1.5)
    t = torch_util.symmetric_csr_to_tensor(H, "cpu", torch.float64)
    assert torch_util.stencil_of(t) == (5, 4, 1.5)
    np.testing.assert_allclose(torch_util.csr_to_scipy(t).toarray(), H.toarray())
    # no CPU path: the product refuses host tensors
    from sttopt_amd import _lib

    with pytest.raises(_lib.DeviceError):
        filters.apply_density_filter(torch.ones(4, 5, dtype=torch.float64), t, torch.ones(20, dtype=torch.float64))

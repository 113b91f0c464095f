"""SIMP compliance under a point load and under staged self-weight, with their
density / time sensitivities (drop-in for sttopt/compliance.py:33-253).

`batched_whole_and_gravity_compliance` -- the entry point of the optimizer step --
runs fused on the device: one prologue kernel builds every row's SIMP density, stage
mask and self-weight load (`sttopt_stage_prologue`), one batched MG-PCG solves all
`1 + nStage` rows (`sttopt_mgpcg`), one epilogue kernel forms the compliances, and
the backward is the hand-derived adjoint of the self-adjoint compliance
(`sttopt_stage_epilogue`: lambda = 2U, the stage-mask chain and the gravity-load term
-2 (C^T U_y)_e), so no adjoint PCG runs at all.  The closed forms are those of the
reference's hand-derived predecessor (tests/reference/compliance.py:24-110).

The single-solve functions (`whole_compliance`, `gravity_compliance`) and the patch
points `_solve_fe` / `_solve_fe_batched` keep the reference's autograd composition
(`FemSolve` + differentiable torch algebra).  When a caller has patched
`_solve_fe_batched` (as a solver-substitution harness does), the batched function
routes through the patched callable instead of the fused path.
"""

from __future__ import annotations

import math

import numpy as np
import torch
from torch import Tensor
from torch.utils.weak import WeakTensorKeyDictionary

from . import _grid, _lib, torch_fem, torch_mg, torch_solve

#: The last fused batched solve: {"n_iter", "row_iters", "row_iter_list", "rows",
#: "levels" (the (nelx, nely) of each multigrid level, finest first)}.
LAST_SOLVE_INFO: dict = {}

#: Coarsening stop of the fused path's hierarchy (None: torch_mg.MAX_COARSE_ELEMENTS).
#: Stage rows whose active material forms small islands floating on Emin need a finer
#: coarsest grid than the default (a coarse space that resolves the islands).
MAX_COARSE_ELEMENTS: int | None = None


def time_mask(tPhys: Tensor, ti: float, beta: float) -> Tensor:
    """Smoothed Heaviside `H_beta(ti - t)`: ~1 where `tPhys < ti`, ~0 after
    (compliance.py:33-46), clamped at 0."""
    a = math.tanh(beta * ti)
    den = a + math.tanh(beta * (1 - ti))
    return torch.clamp(1 - (a + torch.tanh(beta * (tPhys - ti))) / den, min=0)


class _StrainEnergy(torch.autograd.Function):
    """c_e = U_e^T KE U_e; dL/dU = 2 K(g) U with the element weights g (a matvec)."""

    @staticmethod
    def forward(ctx, U, KE, nelx, nely):
        dev = _lib.require_device(U)
        ndof = U.shape[-1]
        batch = tuple(U.shape[:-1])
        rows = max(_grid.numel(batch), 1)
        u2, su = _grid.rows_of(U, ndof, batch, torch.float64)
        ce = torch.empty(rows, nelx * nely, dtype=torch.float64, device=dev)
        d = ndof // ((nelx + 1) * (nely + 1))
        _lib.check(
            _lib.load().sttopt_strain_energy(
                d, nelx, nely, rows, u2.data_ptr(), su, _lib.ke_array(KE), ce.data_ptr(),
                nelx * nely, _lib.stream_handle(dev),
            ),
            "sttopt_strain_energy",
        )
        ctx.save_for_backward(U, KE)
        ctx.grid = (nelx, nely)
        return ce.reshape(*batch, nelx * nely).to(U.dtype)

    @staticmethod
    def backward(ctx, g):
        U, KE = ctx.saved_tensors
        nelx, nely = ctx.grid
        gU = 2 * torch_fem._apply_native(
            _lib.MODE_MATVEC, U, g.to(U.dtype).contiguous(), KE, nelx, nely, U.shape[-1]
        )
        return gU, None, None, None


def _element_strain_energy(U: Tensor, edofMat: Tensor, KE: Tensor, nely: int, nelx: int) -> Tensor:
    """Per-element strain energy reshaped to `(nely, nelx)` (compliance.py:49-59)."""
    _grid.grid_of(edofMat, _grid.dofs_per_node(KE), U.shape[-1])
    return _StrainEnergy.apply(U, KE, nelx, nely).reshape(nely, nelx)


def _solve_fe(KE, xPhys, Emin, Emax, penal, edofMat, freedofs, F, ndof, *, x0=None):
    """Single solve of `K(xPhys) U = F` through `FemSolve` (compliance.py:62-87)."""
    nely, nelx = xPhys.shape
    density = torch_fem.simp_density(xPhys, Emin, Emax, penal)
    mask = torch_fem.free_mask(ndof, freedofs, device=xPhys.device)
    return torch_solve.femsolve(density.flatten(), F, edofMat, KE, mask, nelx, nely, x0=x0)


def _solve_fe_batched(KE, density, edofMat, mask, nelx, nely, F, *, x0=None):
    """Batched solve patch point (compliance.py:90-108)."""
    return torch_solve.femsolve(density, F, edofMat, KE, mask, nelx, nely, x0=x0)


_FUSED_SOLVER = _solve_fe_batched


def whole_compliance(xPhys, KE, edofMat, Emin, Emax, penal, freedofs, F, ndof, *, x0=None):
    """SIMP compliance under a fixed load, `(c, U)` (compliance.py:111-134)."""
    U = _solve_fe(KE, xPhys, Emin, Emax, penal, edofMat, freedofs, F, ndof, x0=x0)
    nely, nelx = xPhys.shape
    ce = _element_strain_energy(U, edofMat, KE, nely, nelx)
    simp = Emin + xPhys**penal * (Emax - Emin)
    return torch.sum(simp * ce), U


def _gravity_load(xPhys, tPhys, ti, C, beta_t, ndof):
    """One stage's `(t_mask, xtJoint, F)`: F_y = -C (x m), F_x = 0 (compliance.py:137-161)."""
    t_mask = time_mask(tPhys, ti, beta_t)
    xtJoint = xPhys * t_mask
    f = -(C @ xtJoint.flatten())
    F = torch.zeros(ndof, dtype=xPhys.dtype, device=xPhys.device)
    F[1::2] = f
    return t_mask, xtJoint, F


def gravity_compliance(xPhys, tPhys, KE, edofMat, Emin, Emax, penal, ti, C, beta_t,
                       freedofs, ndof, *, x0=None):
    """Self-weight compliance of the stage at time `ti`, `(cg, U)` (compliance.py:164-195)."""
    _, xtJoint, F = _gravity_load(xPhys, tPhys, ti, C, beta_t, ndof)
    U = _solve_fe(KE, xtJoint, Emin, Emax, penal, edofMat, freedofs, F, ndof, x0=x0)
    nely, nelx = xPhys.shape
    ce = _element_strain_energy(U, edofMat, KE, nely, nelx)
    simp = Emin + xtJoint**penal * (Emax - Emin)
    return torch.sum(simp * ce), U


# -------------------------------------------------------------- fused path --

_fe_cache = WeakTensorKeyDictionary()
# free-dof masks per freedofs tensor (value: version, ndof, device, mask) and stage-time
# tensors per (times, device): rebuilding them each step copies host data to the
# device, which waits for the stream.  A new mask is complete (its stream synchronized)
# before it is shared, so any stream may read it; the times tensor is copied from host
# memory synchronously.
_mask_cache = WeakTensorKeyDictionary()
_times_cache: dict = {}


def _free_mask_cached(ndof: int, freedofs, device) -> Tensor:
    if not isinstance(freedofs, Tensor):
        return torch_fem.free_mask(ndof, freedofs, device=device)
    hit = _mask_cache.get(freedofs)
    if hit is not None and hit[0] == freedofs._version and hit[1] == ndof and hit[2] == device:
        return hit[3]
    mask = torch_fem.free_mask(ndof, freedofs, device=device)
    if mask.is_cuda:
        torch.cuda.current_stream(mask.device).synchronize()
    _mask_cache[freedofs] = (freedofs._version, ndof, device, mask)
    return mask


def _stage_times_tensor(times: list, dev) -> Tensor:
    key = (tuple(times), str(dev))
    hit = _times_cache.get(key)
    if hit is None:
        if len(_times_cache) > 32:
            _times_cache.clear()
        hit = _times_cache[key] = torch.tensor(times, dtype=torch.float64, device=dev)
    return hit


def gravity_weight(C, nelx: int, nely: int) -> float:
    """`fe / 4`-stencil weight `fe` of a self-weight matrix C.

    The fused kernels apply C as its stencil (each element's `fe / 4` on its four
    corners), so C must have that structure (`gravity.gravity_load_matrix`, possibly
    scaled): a ValueError otherwise.  A float is taken as `fe` itself.
    """
    if isinstance(C, (int, float)):
        return float(C)
    nnodes, nel = (nelx + 1) * (nely + 1), nelx * nely
    if isinstance(C, Tensor):
        hit = _fe_cache.get(C)
        if hit is not None:
            return hit
        if tuple(C.shape) != (nnodes, nel):
            raise ValueError(f"C has shape {tuple(C.shape)}, expected {(nnodes, nel)}")
        Cc = C.to_sparse_coo().coalesce() if C.layout != torch.sparse_coo else C.coalesce()
        idx = Cc.indices()
        vals = Cc.values()
        if vals.numel() != 4 * nel or float(vals.max() - vals.min()) != 0.0:
            raise ValueError("C is not a uniform four-corner self-weight stencil")
        edof = torch_mg.element_dof_map_tensor(nelx, nely, C.device, 1)
        order = torch.argsort(idx[1] * nnodes + idx[0])
        rows_sorted = idx[0][order].reshape(nel, 4).sort(dim=1).values
        if not torch.equal(rows_sorted, edof.sort(dim=1).values):
            raise ValueError("C does not put each element's weight on its own corners")
        fe = 4.0 * float(vals[0])
        _fe_cache[C] = fe
        return fe
    # scipy sparse
    Cc = C.tocoo()
    if Cc.shape != (nnodes, nel) or Cc.nnz != 4 * nel or np.ptp(Cc.data) != 0:
        raise ValueError("C is not a uniform four-corner self-weight stencil")
    return 4.0 * float(Cc.data[0])


class _SpaceTimeCompliance(torch.autograd.Function):
    """Outputs the `1 + nStage` compliances and U (not differentiable).  Backward:
    the closed-form adjoint of `sum_row g_row c_row` w.r.t. xPhys and tPhys."""

    @staticmethod
    def forward(ctx, xPhys, tPhys, KE, Emin, Emax, penal, mask, F, fe, beta_t, times, x0):
        dev = _lib.require_device(xPhys, tPhys, mask, F)
        nely, nelx = xPhys.shape
        nel, ndof = nelx * nely, mask.shape[-1]
        ns = len(times)
        x = xPhys.detach().to(torch.float64).contiguous().reshape(-1)
        t = tPhys.detach().to(torch.float64).contiguous().reshape(-1)
        tk = _stage_times_tensor(times, dev) if ns else None
        fext = F.detach().to(torch.float64).contiguous()
        density = torch.empty(1 + ns, nel, dtype=torch.float64, device=dev)
        load = torch.empty(1 + ns, ndof, dtype=torch.float64, device=dev)
        lib = _lib.load()
        stream = _lib.stream_handle(dev)
        _lib.check(
            lib.sttopt_stage_prologue(
                nelx, nely, ns, x.data_ptr(), t.data_ptr(), _lib.ptr(tk), beta_t, Emin, Emax,
                penal, fe, fext.data_ptr(), density.data_ptr(), load.data_ptr(), stream,
            ),
            "sttopt_stage_prologue",
        )
        edof = torch_mg.element_dof_map_tensor(nelx, nely, dev, 2)
        mce = torch_mg.MAX_COARSE_ELEMENTS if MAX_COARSE_ELEMENTS is None else MAX_COARSE_ELEMENTS
        levels = torch_mg.build_hierarchy(density, edof, KE, mask, nelx, nely,
                                          max_coarse_elements=mce)
        precond = torch_mg.VCycle(levels)
        info: dict = {}
        U, n_iter = torch_mg.native_mgpcg(
            precond, load, rtol=torch_solve.DEFAULT_CG_RTOL, max_iter=500, x0=x0, info=info
        )
        LAST_SOLVE_INFO.clear()
        LAST_SOLVE_INFO.update(n_iter=n_iter, row_iters=info["row_iters"],
                               row_iter_list=info["row_iter_list"], rows=1 + ns,
                               levels=[(lv.nelx, lv.nely) for lv in levels])
        comp = torch.empty(1 + ns, dtype=torch.float64, device=dev)
        scratch = torch.empty(lib.sttopt_stage_scratch_count(nelx, nely, ns),
                              dtype=torch.float64, device=dev)
        ke = KE.detach().to(torch.float64)
        _lib.check(
            lib.sttopt_stage_epilogue(
                nelx, nely, ns, x.data_ptr(), t.data_ptr(), _lib.ptr(tk), beta_t, Emin, Emax,
                penal, fe, _lib.ke_array(ke), U.data_ptr(), density.data_ptr(), None,
                comp.data_ptr(), None, None, scratch.data_ptr(), stream,
            ),
            "sttopt_stage_epilogue",
        )
        ctx.save_for_backward(x, t, U)
        ctx.consts = (tk, beta_t, Emin, Emax, penal, fe, ke, nelx, nely, ns)
        ctx.mark_non_differentiable(U)
        # U's gradient is never used: no rows x ndof zero tensor for it in the backward
        ctx.set_materialize_grads(False)
        return comp.to(xPhys.dtype), U

    @staticmethod
    def backward(ctx, g_comp, g_U):
        x, t, U = ctx.saved_tensors
        tk, beta_t, Emin, Emax, penal, fe, ke, nelx, nely, ns = ctx.consts
        dev = x.device
        if g_comp is None:
            return (None,) * 12
        g = g_comp.detach().to(torch.float64).contiguous()
        dx = torch.empty(nelx * nely, dtype=torch.float64, device=dev)
        dt = torch.empty(nelx * nely, dtype=torch.float64, device=dev)
        _lib.check(
            _lib.load().sttopt_stage_epilogue(
                nelx, nely, ns, x.data_ptr(), t.data_ptr(), _lib.ptr(tk), beta_t, Emin, Emax,
                penal, fe, _lib.ke_array(ke), U.data_ptr(), None, g.data_ptr(), None,
                dx.data_ptr(), dt.data_ptr(), None, _lib.stream_handle(dev),
            ),
            "sttopt_stage_epilogue",
        )
        return (dx.reshape(nely, nelx), dt.reshape(nely, nelx)) + (None,) * 10


def batched_whole_and_gravity_compliance(
    xPhys: Tensor,
    tPhys: Tensor,
    KE: Tensor,
    edofMat: Tensor,
    Emin: float,
    Emax: float,
    penal: float,
    freedofs: Tensor,
    F: Tensor,
    ndof: int,
    C,
    beta_t: float,
    stage_times: list[float],
    *,
    x0: Tensor | None = None,
) -> tuple[Tensor, list[Tensor], Tensor]:
    """Whole-structure compliance plus every stage's self-weight compliance from one
    batched solve, `(c, [cg, ...], U)` (compliance.py:198-253)."""
    _lib.require_device(xPhys, tPhys, F)
    nely, nelx = xPhys.shape
    mask = _free_mask_cached(ndof, freedofs, xPhys.device)
    if _solve_fe_batched is not _FUSED_SOLVER or KE.shape[-1] != 8:
        return _batched_composed(xPhys, tPhys, KE, edofMat, Emin, Emax, penal, mask, F, ndof,
                                 C, beta_t, stage_times, x0)
    _grid.grid_of(edofMat, 2, ndof)
    fe = gravity_weight(C, nelx, nely)
    comp, U = _SpaceTimeCompliance.apply(
        xPhys, tPhys, KE, Emin, Emax, penal, mask, F, fe, beta_t,
        [float(v) for v in stage_times], None if x0 is None else x0.detach(),
    )
    # one unbind: its backward stacks the scalar gradients in one kernel, where one
    # indexing view per stage would zero-fill and scatter one small tensor each
    parts = comp.unbind(0)
    return parts[0], list(parts[1:]), U


def _batched_composed(xPhys, tPhys, KE, edofMat, Emin, Emax, penal, mask, F, ndof, C,
                      beta_t, stage_times, x0):
    """The reference's composition: torch algebra around the (possibly patched)
    `_solve_fe_batched`."""
    nely, nelx = xPhys.shape
    density_rows = [torch_fem.simp_density(xPhys, Emin, Emax, penal).flatten()]
    F_rows = [F]
    xts = []
    for ti in stage_times:
        _, xt, Fg = _gravity_load(xPhys, tPhys, ti, C, beta_t, ndof)
        density_rows.append(torch_fem.simp_density(xt, Emin, Emax, penal).flatten())
        F_rows.append(Fg)
        xts.append(xt)
    U = _solve_fe_batched(KE, torch.stack(density_rows), edofMat, mask, nelx, nely,
                          torch.stack(F_rows), x0=x0)
    ce = _element_strain_energy(U[0], edofMat, KE, nely, nelx)
    c = torch.sum((Emin + xPhys**penal * (Emax - Emin)) * ce)
    stages = []
    for i, xt in enumerate(xts):
        ce = _element_strain_energy(U[1 + i], edofMat, KE, nely, nelx)
        stages.append(torch.sum((Emin + xt**penal * (Emax - Emin)) * ce))
    return c, stages, U

"""The multigrid-CG solve as an autograd Function, on the gfx950 HIP kernels.

Drop-in for the reference's `sttopt.torch_solve` (sttopt/torch_solve.py:49-287): the
16-argument `FemSolve.apply`, `femsolve`, `lifted_femsolve` and `DEFAULT_CG_RTOL`.

For `K U = F` with `K = sum_e d_e KE` and a scalar `L(U)`, `g = dL/dU`:
`lambda = K^-1 g`, `dL/dF = lambda`, `dL/dd_e = -(lambda_e KE) . U_e`,
`dL/dreaction_e = -sum(lambda_e * U_e) / 4`.  The forward's hierarchy is kept on
`ctx` and reused by the adjoint solve, which warm-starts from `alpha U`,
`alpha = (U . g) / (U . F)` -- exact for compliance-type objectives, where the
adjoint then returns at iteration 0.  Both solves run in `sttopt_mgpcg`; the element
terms of the gradient in `sttopt_adjoint_grad`.
"""

from __future__ import annotations

import torch
from torch import Tensor

from . import _grid, _lib, torch_fem, torch_mg

#: Default CG relative-residual tolerance (torch_solve.py:45-49).
DEFAULT_CG_RTOL = 1e-8


def _element_grads(U: Tensor, lam: Tensor, KE: Tensor, nelx: int, nely: int, ndof: int,
                   want_density: bool, want_reaction: bool):
    dev = _lib.require_device(U, lam)
    batch = _grid.batch_shape(U.shape[:-1], lam.shape[:-1])
    rows = max(_grid.numel(batch), 1)
    u2, su = _grid.rows_of(U, ndof, batch, torch.float64)
    l2, sl = _grid.rows_of(lam, ndof, batch, torch.float64)
    nel = nelx * nely
    gd = torch.empty(rows, nel, dtype=torch.float64, device=dev) if want_density else None
    gr = torch.empty(rows, nel, dtype=torch.float64, device=dev) if want_reaction else None
    d = ndof // ((nelx + 1) * (nely + 1))
    _lib.check(
        _lib.load().sttopt_adjoint_grad(
            d, nelx, nely, rows, u2.data_ptr(), su, l2.data_ptr(), sl, _lib.ke_array(KE),
            _lib.ptr(gd), nel, _lib.ptr(gr), nel, _lib.stream_handle(dev),
        ),
        "sttopt_adjoint_grad",
    )
    shape = (*batch, nel)
    return (None if gd is None else gd.reshape(shape)), (None if gr is None else gr.reshape(shape))


class FemSolve(torch.autograd.Function):
    """`K @ U = F` as a differentiable op (torch_solve.py:52-173).  A
    `CGConvergenceError` of the forward or the adjoint solve propagates unchanged."""

    @staticmethod
    def forward(ctx, density, F, reaction, edofMat, KE, mask, nelx, nely, rtol, max_iter,
                x0, omega, n_smooth, gamma, max_coarse_elements, info):
        levels = torch_mg.build_hierarchy(
            density, edofMat, KE, mask, nelx, nely, reaction=reaction,
            max_coarse_elements=max_coarse_elements,
        )
        precond = torch_mg.VCycle(levels, omega=omega, n_smooth=n_smooth, gamma=gamma)
        b = torch_fem.project(F, mask)
        U, n_iter = torch_fem.pcg(levels[0], b, precond, rtol=rtol, max_iter=max_iter, x0=x0)
        U = torch_fem.project(U, mask)
        ctx.save_for_backward(U, density, F)
        ctx.levels = levels
        ctx.precond = precond
        ctx.KE = KE
        ctx.mask = mask
        ctx.grid = (nelx, nely, mask.shape[-1])
        ctx.rtol, ctx.max_iter = rtol, max_iter
        ctx.info = info
        if info is not None:
            info["forward_n_iter"] = n_iter
        return U

    @staticmethod
    def backward(ctx, grad_output):
        U, density, F = ctx.saved_tensors
        g = grad_output.to(U.dtype)
        UF = (U * F).sum(dim=-1)
        Ug = (U * g).sum(dim=-1)
        alpha = torch_fem.safe_div(Ug, UF)
        x0 = alpha[..., None] * U
        b = torch_fem.project(g, ctx.mask)
        lam, n_iter = torch_fem.pcg(
            ctx.levels[0], b, ctx.precond, rtol=ctx.rtol, max_iter=ctx.max_iter, x0=x0
        )
        lam = torch_fem.project(lam, ctx.mask)
        if ctx.info is not None:
            ctx.info["backward_n_iter"] = n_iter
        want_d, want_r = ctx.needs_input_grad[0], ctx.needs_input_grad[2]
        grad_density = grad_reaction = None
        if want_d or want_r:
            nelx, nely, ndof = ctx.grid
            grad_density, grad_reaction = _element_grads(
                U, lam, ctx.KE, nelx, nely, ndof, want_d, want_r
            )
            if grad_density is not None:
                grad_density = grad_density.to(density.dtype)
        grad_F = lam if ctx.needs_input_grad[1] else None
        return (grad_density, grad_F, grad_reaction) + (None,) * 13


def femsolve(
    density: Tensor,
    F: Tensor,
    edofMat: Tensor,
    KE: Tensor,
    mask: Tensor,
    nelx: int,
    nely: int,
    *,
    rtol: float = DEFAULT_CG_RTOL,
    max_iter: int = 500,
    x0: Tensor | None = None,
    omega: float = 0.6,
    n_smooth: int = 2,
    gamma: int = 1,
    max_coarse_elements: int = torch_mg.MAX_COARSE_ELEMENTS,
    info: dict | None = None,
    reaction: Tensor | None = None,
) -> Tensor:
    """`FemSolve.apply` with the keyword defaults of the reference (torch_solve.py:176-243).
    `x0` must already be detached."""
    assert x0 is None or not x0.requires_grad, (
        "x0 must already be detached -- FemSolve.backward returns None for it "
        "unconditionally, so no caller should be routing gradient through it."
    )
    _lib.require_device(density, F, mask, reaction, x0)
    return FemSolve.apply(density, F, reaction, edofMat, KE, mask, nelx, nely, rtol,
                          max_iter, x0, omega, n_smooth, gamma, max_coarse_elements, info)


def lifted_femsolve(
    density: Tensor,
    F: Tensor,
    g: Tensor,
    edofMat: Tensor,
    KE: Tensor,
    mask: Tensor,
    nelx: int,
    nely: int,
    *,
    reaction: Tensor | None = None,
    **solve_kwargs,
) -> Tensor:
    """`femsolve` with Dirichlet values `g` on the fixed dofs, by lifting:
    `u = femsolve(F - K g_fixed) + g_fixed` (torch_solve.py:246-287)."""
    ndof = mask.shape[-1]
    lift = g * ~mask
    K_lift = torch_fem.matvec(lift, density, edofMat, KE, ndof)
    if reaction is not None:
        K_lift = K_lift + torch_fem.lumped_reaction(reaction, edofMat, ndof) * lift
    free = femsolve(density, F - K_lift, edofMat, KE, mask, nelx, nely, reaction=reaction,
                    **solve_kwargs)
    return free + lift

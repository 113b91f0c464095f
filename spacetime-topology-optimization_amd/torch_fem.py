"""Matrix-free Q4 stiffness operator and batched PCG, on the gfx950 HIP kernels.

Drop-in for the reference's `sttopt.torch_fem` (sttopt/torch_fem.py:30-383): same
names, arguments, broadcasting, return values and errors.  The operator apply, its
diagonal and the lumped reaction run in `libsttopt_hip.so`; the element dof map is
validated once and then never read by a kernel (indices follow from the structured
grid).  Boundary conditions are projections, `A(v) = P(K P v + r v)`.

`pcg` keeps the reference's generic contract for any pair of callables (row
retirement through `select`, `COMPACTION_RATIO`, `CGConvergenceError` with the
caller's row indices).  When `apply_A` is level 0 of a `torch_mg` hierarchy and
`apply_M` its `VCycle`, the whole iteration instead runs on the device in
`sttopt_mgpcg` (per-row convergence flags, fp64 reductions, no per-iteration host
sync).

All kernel-backed functions need tensors on a ROCm device and raise
`DeviceError` otherwise; there is no CPU path.
"""

from __future__ import annotations

import torch
from torch import Tensor

from . import _grid, _lib

#: Compact the batch of a generic `pcg` call once the active rows have fallen to
#: this fraction of their count at the last compaction (sttopt/torch_fem.py:30).
COMPACTION_RATIO = 0.75


class CGConvergenceError(RuntimeError):
    """PCG did not reach `rtol` within `max_iter` (sttopt/torch_fem.py:33-52).

    :ivar residuals: final relative residual of every row of the caller's batch
    :ivar n_iter: iterations run (`max_iter`)
    :ivar rtol: the tolerance asked for
    """

    def __init__(self, residuals: Tensor, n_iter: int, rtol: float) -> None:
        self.residuals = residuals
        self.n_iter = n_iter
        self.rtol = rtol
        worst = residuals.max().item()
        failed = torch.nonzero(~(residuals <= rtol)).flatten().tolist()
        super().__init__(
            f"CG did not converge to rtol={rtol:g} within {n_iter} iterations "
            f"(worst relative residual {worst:g}); failed batch indices {failed}"
        )


def simp_density(xPhys: Tensor, Emin: float, Emax: float, penal: float) -> Tensor:
    """SIMP stiffness scale `Emin + xPhys**penal * (Emax - Emin)` (torch_fem.py:55-59)."""
    return Emin + xPhys**penal * (Emax - Emin)


def free_mask(ndof: int, freedofs: Tensor, device=None) -> Tensor:
    """Bool mask, True at the free dofs (torch_fem.py:101-107)."""
    mask = torch.zeros(ndof, dtype=torch.bool, device=device)
    mask[freedofs.to(mask.device)] = True
    return mask


def project(v: Tensor, mask: Tensor) -> Tensor:
    """Zero the fixed dofs of `v` (torch_fem.py:110-114)."""
    return v * mask


def safe_div(num: Tensor, den: Tensor) -> Tensor:
    """`num / den`, 0 where `den == 0` (torch_fem.py:184-192)."""
    ok = den != 0
    return torch.where(ok, num / torch.where(ok, den, torch.ones_like(den)), torch.zeros_like(den))


# ----------------------------------------------------------------- operator --


def _apply_native(
    mode: int,
    U: Tensor,
    density: Tensor,
    KE: Tensor,
    nelx: int,
    nely: int,
    ndof: int,
    mask: Tensor | None = None,
    rdiag: Tensor | None = None,
) -> Tensor:
    dev = _lib.require_device(U, density, mask, rdiag)
    dtype = U.dtype
    if density.dtype != dtype:
        raise TypeError(f"U is {dtype} but density is {density.dtype}")
    nel = nelx * nely
    shapes = [U.shape[:-1], density.shape[:-1]]
    if rdiag is not None:
        shapes.append(rdiag.shape[:-1])
    batch = _grid.batch_shape(*shapes)
    rows = _grid.numel(batch)
    u2, su = _grid.rows_of(U, ndof, batch)
    d2, sd = _grid.rows_of(density, nel, batch)
    r2, sr = (None, 0) if rdiag is None else _grid.rows_of(rdiag, ndof, batch, dtype)
    out = torch.empty(rows, ndof, dtype=dtype, device=dev)
    bits = _grid.mask_bits(mask) if mask is not None else None
    d = ndof // ((nelx + 1) * (nely + 1))
    _lib.check(
        _lib.load().sttopt_apply(
            _lib.dtype_code(dtype), mode, d, nelx, nely, rows, u2.data_ptr(), su,
            d2.data_ptr(), sd, _lib.ke_array(KE), _lib.ptr(bits), _lib.ptr(r2), sr,
            out.data_ptr(), ndof, _lib.stream_handle(dev),
        ),
        "sttopt_apply",
    )
    return out.reshape(*batch, ndof)


def _element_contract(U: Tensor, lam: Tensor, KE: Tensor, nelx: int, nely: int, ndof: int) -> Tensor:
    """`(lam_e KE) . U_e` per element and row, on the device (sttopt_adjoint_grad)."""
    dev = _lib.require_device(U, lam)
    batch = _grid.batch_shape(U.shape[:-1], lam.shape[:-1])
    rows = _grid.numel(batch)
    u2, su = _grid.rows_of(U, ndof, batch, torch.float64)
    l2, sl = _grid.rows_of(lam, ndof, batch, torch.float64)
    out = torch.empty(rows, nelx * nely, dtype=torch.float64, device=dev)
    d = ndof // ((nelx + 1) * (nely + 1))
    _lib.check(
        _lib.load().sttopt_adjoint_grad(
            d, nelx, nely, rows, u2.data_ptr(), su, l2.data_ptr(), sl, _lib.ke_array(KE),
            out.data_ptr(), nelx * nely, None, 0, _lib.stream_handle(dev),
        ),
        "sttopt_adjoint_grad",
    )
    return (-out).reshape(*batch, nelx * nely)


class _Matvec(torch.autograd.Function):
    """`K(density) U` with its adjoint: dL/dU = K g (K symmetric), dL/d density_e =
    (g_e KE) . U_e -- both on the HIP kernels."""

    @staticmethod
    def forward(ctx, U, density, KE, nelx, nely, ndof):
        ctx.save_for_backward(U, density, KE)
        ctx.grid = (nelx, nely, ndof)
        return _apply_native(_lib.MODE_MATVEC, U, density, KE, nelx, nely, ndof)

    @staticmethod
    def backward(ctx, g):
        U, density, KE = ctx.saved_tensors
        nelx, nely, ndof = ctx.grid
        g = g.contiguous()
        gU = gd = None
        if ctx.needs_input_grad[0]:
            gU = _apply_native(_lib.MODE_MATVEC, g, density.to(g.dtype), KE, nelx, nely, ndof)
        if ctx.needs_input_grad[1]:
            gd = _element_contract(U, g, KE, nelx, nely, ndof).to(density.dtype)
        return gU, gd, None, None, None, None


def matvec(U: Tensor, density: Tensor, edofMat: Tensor, KE: Tensor, ndof: int) -> Tensor:
    """Matrix-free `K @ U` with `K = sum_e density_e KE_e` (torch_fem.py:62-83).

    Batch dims of `U` and `density` broadcast against each other.  Differentiable
    w.r.t. `U` and `density`.
    """
    _lib.require_device(U, density)
    d = _grid.dofs_per_node(KE)
    nelx, nely = _grid.grid_of(edofMat, d, ndof)
    if torch.is_grad_enabled() and (U.requires_grad or density.requires_grad):
        return _Matvec.apply(U, density, KE, nelx, nely, ndof)
    return _apply_native(_lib.MODE_MATVEC, U, density, KE, nelx, nely, ndof)


def matvec_diagonal(density: Tensor, edofMat: Tensor, KE: Tensor, ndof: int) -> Tensor:
    """`diag(K)` (torch_fem.py:86-98)."""
    return _diagonal(density, edofMat, KE, ndof, None, None)


def _diagonal(density, edofMat, KE, ndof, mask, reaction_diag) -> Tensor:
    dev = _lib.require_device(density, mask, reaction_diag)
    d = _grid.dofs_per_node(KE)
    nelx, nely = _grid.grid_of(edofMat, d, ndof)
    den = density.detach().to(torch.float64)
    shapes = [den.shape[:-1]] + ([reaction_diag.shape[:-1]] if reaction_diag is not None else [])
    batch = _grid.batch_shape(*shapes)
    rows = _grid.numel(batch)
    d2, sd = _grid.rows_of(den, nelx * nely, batch)
    r2, sr = (None, 0)
    if reaction_diag is not None:
        r2, sr = _grid.rows_of(reaction_diag.detach(), ndof, batch, torch.float64)
    out = torch.empty(rows, ndof, dtype=torch.float64, device=dev)
    bits = _grid.mask_bits(mask) if mask is not None else None
    _lib.check(
        _lib.load().sttopt_diagonal(
            _lib.F64, d, nelx, nely, rows, d2.data_ptr(), sd, _lib.ke_array(KE),
            _lib.ptr(bits), _lib.ptr(r2), sr, out.data_ptr(), ndof, _lib.stream_handle(dev),
        ),
        "sttopt_diagonal",
    )
    return out.reshape(*batch, ndof).to(density.dtype)


def lumped_reaction(reaction: Tensor, edofMat: Tensor, ndof: int) -> Tensor:
    """Diagonal of a lumped reaction term: each element adds `reaction_e / 4` to each of
    its dofs (torch_fem.py:117-132).  Differentiable w.r.t. `reaction`."""
    k = edofMat.shape[-1]
    d = k // 4
    nelx, nely = _grid.grid_of(edofMat, d, ndof)
    if torch.is_grad_enabled() and reaction.requires_grad:
        return _Lumped.apply(reaction, nelx, nely, d, ndof, edofMat)
    return _lumped_native(reaction, nelx, nely, d, ndof)


def _lumped_native(reaction, nelx, nely, d, ndof) -> Tensor:
    dev = _lib.require_device(reaction)
    batch = tuple(reaction.shape[:-1])
    rows = _grid.numel(batch)
    r2, sr = _grid.rows_of(reaction, nelx * nely, batch)
    out = torch.empty(rows, ndof, dtype=reaction.dtype, device=dev)
    _lib.check(
        _lib.load().sttopt_lumped_reaction(
            _lib.dtype_code(reaction.dtype), d, nelx, nely, rows, r2.data_ptr(), sr,
            out.data_ptr(), ndof, _lib.stream_handle(dev),
        ),
        "sttopt_lumped_reaction",
    )
    return out.reshape(*batch, ndof)


class _Lumped(torch.autograd.Function):
    @staticmethod
    def forward(ctx, reaction, nelx, nely, d, ndof, edofMat):
        ctx.edofMat = edofMat
        return _lumped_native(reaction, nelx, nely, d, ndof)

    @staticmethod
    def backward(ctx, g):
        # d(out)/d(reaction_e) puts 1/4 on each of the element's dofs
        return g[..., ctx.edofMat].sum(-1) / 4, None, None, None, None, None


def operator(
    v: Tensor,
    density: Tensor,
    edofMat: Tensor,
    KE: Tensor,
    ndof: int,
    mask: Tensor,
    reaction_diag: Tensor | None = None,
) -> Tensor:
    """The projected operator `A(v) = P(K P v + reaction_diag * P v)` (torch_fem.py:135-159)."""
    _lib.require_device(v, density, mask, reaction_diag)
    d = _grid.dofs_per_node(KE)
    nelx, nely = _grid.grid_of(edofMat, d, ndof)
    needs_grad = torch.is_grad_enabled() and any(
        t is not None and t.requires_grad for t in (v, density, reaction_diag)
    )
    if needs_grad:
        pv = project(v, mask)
        Kv = matvec(pv, density, edofMat, KE, ndof)
        if reaction_diag is not None:
            Kv = Kv + reaction_diag * pv
        return project(Kv, mask)
    return _apply_native(
        _lib.MODE_OPERATOR, v, density, KE, nelx, nely, ndof, mask, reaction_diag
    )


def jacobi_preconditioner_diag(
    density: Tensor,
    edofMat: Tensor,
    KE: Tensor,
    ndof: int,
    mask: Tensor,
    reaction_diag: Tensor | None = None,
) -> Tensor:
    """Diagonal of the projected operator with 1 at fixed dofs (torch_fem.py:162-181)."""
    return _diagonal(density, edofMat, KE, ndof, mask, reaction_diag)


# ---------------------------------------------------------------------- PCG --


def _report(info: dict | None, row_iters: int) -> None:
    if info is not None:
        info["row_iters"] = row_iters


def pcg(
    apply_A,
    b: Tensor,
    apply_M,
    *,
    rtol: float = 1e-8,
    max_iter: int = 10000,
    x0: Tensor | None = None,
    info: dict | None = None,
) -> tuple[Tensor, int]:
    """Preconditioned CG on `apply_A(x) = b`, batched over leading dims of `b`
    (torch_fem.py:201-343).

    Every row iterates until `||r|| / ||b|| <= rtol`; rows retire as they converge.
    Raises `CGConvergenceError` (never returns an unconverged result) when a row is
    still above `rtol` after `max_iter` iterations.  `apply_M` must be a fixed SPD
    operator.

    :return: `(x, n_iter)`, `n_iter` the iteration count of the slowest row
    """
    _lib.require_device(b)
    native = getattr(apply_M, "_native_pcg", None)
    if native is not None:
        out = native(apply_A, b, rtol=rtol, max_iter=max_iter, x0=x0, info=info)
        if out is not None:
            return out
    return _generic_pcg(apply_A, b, apply_M, rtol, max_iter, x0, info)


def _generic_pcg(apply_A, b, apply_M, rtol, max_iter, x0, info):
    """PCG over arbitrary callables, with the reference's compaction schedule: the
    converged rows are banked and the operators narrowed through `select` once the
    active rows fall to `COMPACTION_RATIO` of the last compaction."""
    bnorm = b.norm(dim=-1)
    bsafe = torch.where(bnorm > 0, bnorm, torch.ones_like(bnorm))
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    r = b - apply_A(x)
    rel = r.norm(dim=-1) / bsafe
    converged = rel <= rtol
    remaining = int(torch.count_nonzero(~converged))
    rows_now = b.shape[0] if b.ndim == 2 else 1
    row_iters = 0
    if remaining == 0:
        _report(info, row_iters)
        return x, 0

    selectable = (
        b.ndim == 2 and hasattr(apply_A, "select") and hasattr(apply_M, "select")
    )
    if selectable:
        x_done = torch.zeros_like(b)
        rel_done = torch.zeros_like(rel)
        active = torch.arange(b.shape[0], device=b.device)
        last_count = b.shape[0]

    def compact(vectors):
        nonlocal apply_A, apply_M, active, last_count, bsafe, rel, rows_now
        done = torch.nonzero(converged).flatten()
        keep = torch.nonzero(~converged).flatten()
        x_done[active[done]] = x[done]
        rel_done[active[done]] = rel[done]
        active = active[keep]
        last_count = keep.numel()
        rows_now = keep.numel()
        apply_A, apply_M = apply_A.select(keep), apply_M.select(keep)
        bsafe, rel = bsafe[keep], rel[keep]
        return [t[keep] for t in vectors]

    def due() -> bool:
        return selectable and remaining <= COMPACTION_RATIO * last_count

    if due():
        x, r = compact([x, r])
    z = apply_M(r)
    p = z.clone()
    rz = (r * z).sum(dim=-1)
    n_iter = 0
    for n_iter in range(1, max_iter + 1):
        row_iters += rows_now
        Ap = apply_A(p)
        alpha = safe_div(rz, (p * Ap).sum(dim=-1))[..., None]
        x = x + alpha * p
        r = r - alpha * Ap
        rel = r.norm(dim=-1) / bsafe
        converged = rel <= rtol
        remaining = int(torch.count_nonzero(~converged))
        if remaining == 0:
            break
        z = apply_M(r)
        rz_new = (r * z).sum(dim=-1)
        p = z + safe_div(rz_new, rz)[..., None] * p
        rz = rz_new
        if due():
            x, r, p, rz = compact([x, r, p, rz])
    else:
        if selectable:
            rel_done[active] = rel
            rel = rel_done
        _report(info, row_iters)
        raise CGConvergenceError(rel, max_iter, rtol)
    _report(info, row_iters)
    if selectable:
        x_done[active] = x
        return x_done, n_iter
    return x, n_iter


def solve(
    F: Tensor,
    xPhys: Tensor,
    edofMat: Tensor,
    KE: Tensor,
    Emin: float,
    Emax: float,
    penal: float,
    mask: Tensor,
    *,
    rtol: float = 1e-8,
    max_iter: int = 10000,
    x0: Tensor | None = None,
) -> tuple[Tensor, int]:
    """SIMP density from `xPhys`, then Jacobi-PCG for `K U = F` (torch_fem.py:346-383).
    Fixed dofs of `F` are ignored and come back exactly 0 in `U`."""
    _lib.require_device(F, xPhys, mask)
    ndof = mask.shape[-1]
    density = simp_density(xPhys, Emin, Emax, penal)

    def apply_A(v):
        return operator(v, density, edofMat, KE, ndof, mask)

    diag = jacobi_preconditioner_diag(density, edofMat, KE, ndof, mask)
    U, n_iter = pcg(
        apply_A, project(F, mask), lambda r: r / diag, rtol=rtol, max_iter=max_iter, x0=x0
    )
    return project(U, mask), n_iter

"""Batched solves of independent instances of uneven size.

The reference batches rows over one grid only (one `edofMat` per call,
sttopt/torch_fem.py:62-83); instances of different sizes are solved one call at a
time.  Here the instances are grouped by grid (and free-dof mask); each group gets its
own multigrid hierarchy, built batched over the group's rows (one row per instance and
load case), and all groups are solved by ONE batched MG-PCG (`sttopt_mgpcg_batch`):
every pass of an iteration is a single launch over the work of all groups (group,
level, tile), the V-cycle visits level l of every group that has it in one launch and
the coarsest solves of all groups together, and rows retire on their own.  The
hierarchies' smoother schedules come from the autotuner (`torch_mg.autotune`).

Differentiable w.r.t. every instance's density and loads: the backward is a second
batched solve for the adjoints, warm-started at alpha U per row as `FemSolve` does
(sttopt/torch_solve.py:117-152), then the element terms -(lambda_e KE) . U_e.

Groups outside the batch solver's scope (an fp64 cycle, or a hierarchy not coarsened
2 x 2 throughout) are solved one `femsolve` per group instead (the same solution).
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict
from dataclasses import dataclass

import torch
from torch import Tensor

from . import _grid, _lib, torch_fem, torch_mg, torch_solve


@dataclass
class Instance:
    """One elasticity instance: grid, SIMP density `(nel,)` (requires_grad allowed),
    load cases `(n_cases, ndof)`, free-dof mask, dof map and KE on the device."""

    nelx: int
    nely: int
    density: Tensor
    loads: Tensor
    mask: Tensor
    edofMat: Tensor
    KE: Tensor


class BatchGroup(ctypes.Structure):
    """`sttopt_batch_group` (include/sttopt_hip.h)."""

    _fields_ = [
        ("mg", ctypes.c_void_p),
        ("b", ctypes.c_void_p),
        ("b_stride", ctypes.c_int64),
        ("x", ctypes.c_void_p),
        ("x_stride", ctypes.c_int64),
        ("work", ctypes.c_void_p),
        ("iters", ctypes.c_void_p),
        ("resid", ctypes.c_void_p),
    ]


def batch_solve(vcycles: list, rhs: list[Tensor], *, rtol: float, max_iter: int,
                x0: list[Tensor] | None = None, info: dict | None = None) -> list[Tensor]:
    """`A_g x_g = b_g` for every group g at once (`sttopt_mgpcg_batch`): `vcycles[g]` the
    group's fp32 `torch_mg.VCycle`, `rhs[g]` its `(rows, ndof)` fp64 right-hand sides
    (fixed dofs ignored), `x0[g]` an optional warm start.

    :return: the solutions, fixed dofs 0
    :raises CGConvergenceError: rows above `rtol` after `max_iter` (group-major order)
    :raises NativeLibraryError: a hierarchy outside the batch solver's scope
    """
    dev = _lib.require_device(*rhs)
    n = len(vcycles)
    groups = (BatchGroup * n)()
    keep, xs, iters, resids = [], [], [], []  # keep: what the native call reads
    for i, (vc, b) in enumerate(zip(vcycles, rhs)):
        lv0 = vc.levels[0]
        rows = b.shape[0]
        mg, _, work = vc._native(rows)
        b2 = b.detach().to(torch.float64).contiguous()
        if x0 is None:
            x = torch.empty(rows, lv0.ndof, dtype=torch.float64, device=dev)
        else:
            x = x0[i].detach().to(torch.float64).contiguous().clone()
        it = torch.zeros(rows, dtype=torch.int32, device=dev)
        rs = torch.zeros(rows, dtype=torch.float64, device=dev)
        keep += [mg, b2, work]
        xs.append(x)
        iters.append(it)
        resids.append(rs)
        groups[i].mg = ctypes.addressof(mg)
        groups[i].b, groups[i].b_stride = b2.data_ptr(), lv0.ndof
        groups[i].x, groups[i].x_stride = x.data_ptr(), lv0.ndof
        groups[i].work = work.data_ptr()
        groups[i].iters, groups[i].resid = it.data_ptr(), rs.data_ptr()
    n_iter = ctypes.c_int32(0)
    status = _lib.load().sttopt_mgpcg_batch(
        ctypes.addressof(groups), n, float(rtol), int(max_iter), torch_mg.CHECK_EVERY,
        1 if x0 is None else 0, ctypes.byref(n_iter), _lib.stream_handle(dev))
    if info is not None:
        flat = torch.cat(iters).tolist()
        info["n_iter"] = int(n_iter.value)
        info["row_iter_list"] = flat
        info["row_iters"] = int(sum(flat))
    if status == _lib.ENOTSPD:
        raise torch.linalg.LinAlgError(torch_mg._NOT_SPD)
    for vc in vcycles:
        vc.levels[-1]._cinv_status = None  # checked by the solve
    if status == _lib.ENOCONV:
        raise torch_fem.CGConvergenceError(torch.cat(resids), int(max_iter), rtol)
    _lib.check(status, "sttopt_mgpcg_batch")
    return xs


class _BatchSolve(torch.autograd.Function):
    """Outputs the solutions of every group; backward: the batched adjoint solve."""

    @staticmethod
    def forward(ctx, spec, *tensors):
        groups, rtol, max_iter, info = spec
        n = len(groups)
        dens, loads = tensors[:n], tensors[n:]
        vcs, rhs = [], []
        for (nelx, nely, edof, KE, mask), d, F in zip(groups, dens, loads):
            levels = torch_mg.build_hierarchy(d.detach(), edof, KE, mask, nelx, nely)
            vcs.append(torch_mg.VCycle(levels, dtype=torch.float32))
            rhs.append(torch_fem.project(F.detach().to(torch.float64), mask))
        fwd: dict = {}
        Us = batch_solve(vcs, rhs, rtol=rtol, max_iter=max_iter, info=fwd)
        Us = [torch_fem.project(U, g[4]) for U, g in zip(Us, groups)]
        if info is not None:
            info.update(forward_n_iter=fwd["n_iter"], forward_row_iters=fwd["row_iters"])
        ctx.spec = spec
        ctx.vcs = vcs
        ctx.save_for_backward(*Us, *loads)
        return tuple(Us)

    @staticmethod
    def backward(ctx, *grads):
        groups, rtol, max_iter, info = ctx.spec
        n = len(groups)
        saved = ctx.saved_tensors
        Us, loads = saved[:n], saved[n:]
        gs = [torch.zeros_like(U) if g is None else g.to(torch.float64) for U, g in zip(Us, grads)]
        x0, rhs = [], []
        for U, F, g, grp in zip(Us, loads, gs, groups):
            alpha = torch_fem.safe_div((U * g).sum(-1), (U * F.to(torch.float64)).sum(-1))
            x0.append(alpha[:, None] * U)
            rhs.append(torch_fem.project(g, grp[4]))
        bwd: dict = {}
        lams = batch_solve(ctx.vcs, rhs, rtol=rtol, max_iter=max_iter, x0=x0, info=bwd)
        if info is not None:
            info["backward_n_iter"] = bwd["n_iter"]
        grad_d, grad_F = [], []
        for i, ((nelx, nely, edof, KE, mask), U, lam) in enumerate(zip(groups, Us, lams)):
            lam = torch_fem.project(lam, mask)
            gd = None
            if ctx.needs_input_grad[1 + i]:
                gd, _ = torch_solve._element_grads(U, lam, KE, nelx, nely, U.shape[-1], True,
                                                   False)
            grad_d.append(gd)
            grad_F.append(lam if ctx.needs_input_grad[1 + n + i] else None)
        return (None, *grad_d, *grad_F)


def _grouped(instances: list[Instance]):
    """Instance indices by (grid, free-dof mask), in first-seen order.  The masks of one
    grid are compared on the device in one pass (one host read per grid)."""
    by_grid: "OrderedDict[tuple, list[int]]" = OrderedDict()
    for idx, inst in enumerate(instances):
        by_grid.setdefault((inst.nelx, inst.nely), []).append(idx)
    out = []
    for (nelx, nely), idxs in by_grid.items():
        masks = [instances[i].mask for i in idxs]
        if all(m is masks[0] for m in masks):
            out.append(((nelx, nely, 0), idxs))
            continue
        stacked = torch.stack(masks)
        # a fingerprint per mask (a fixed random weighting of its free dofs) proposes the
        # groups; one exact comparison with each group's first mask confirms them
        g = torch.Generator(device="cpu").manual_seed(1)
        w = torch.rand(stacked.shape[1], generator=g, dtype=torch.float64).to(stacked.device)
        prints = (stacked.to(torch.float64) * w).sum(-1).tolist()
        rep: dict = {}
        first = [rep.setdefault(fp, j) for j, fp in enumerate(prints)]
        same = (stacked == stacked[torch.tensor(first, device=stacked.device)]).all(-1)
        if not bool(same.all()):
            first = list(range(len(masks)))  # a fingerprint collision: no sharing at all
        sub: "OrderedDict[int, list[int]]" = OrderedDict()
        for j, f in enumerate(first):
            sub.setdefault(f, []).append(idxs[j])
        out += [((nelx, nely, f), members) for f, members in sub.items()]
    return out


def solve_instances(instances: list[Instance], *, rtol: float = 1e-8, max_iter: int = 500,
                    info: dict | None = None) -> list[Tensor]:
    """`U` of every instance, shape `(n_cases, ndof)`, differentiable w.r.t. each
    instance's density and loads.

    :param info: receives `{"groups": [(nelx, nely, rows), ...], "batched": bool,
        "forward_n_iter", "forward_row_iters", "backward_n_iter" (after backward)}`
    """
    if not instances:
        return []
    _lib.require_device(*(i.density for i in instances))
    items = _grouped(instances)
    groups, dens, loads, cases = [], [], [], []
    for key, idxs in items:
        first = instances[idxs[0]]
        ncases = [instances[i].loads.shape[0] for i in idxs]
        # each instance's density once per load case of its own
        dens.append(torch.cat([instances[i].density.reshape(1, -1).expand(nc, -1)
                               for i, nc in zip(idxs, ncases)]))
        loads.append(torch.cat([instances[i].loads for i in idxs]))
        _grid.grid_of(first.edofMat, 2 if first.KE.shape[-1] == 8 else 1, first.loads.shape[-1])
        groups.append((first.nelx, first.nely, first.edofMat, first.KE, first.mask))
        cases.append(ncases)
    if info is not None:
        info["groups"] = [(g[0], g[1], int(d.shape[0])) for g, d in zip(groups, dens)]
    batched = torch_mg.VCYCLE_DTYPE == torch.float32
    for (nelx, nely, _, _, _) in groups:
        factors = _coarsening(nelx, nely)
        batched = batched and len(factors) >= 1 and all(k == (2, 2) for k in factors)
    if info is not None:
        info["batched"] = batched
    if batched:
        Us = _BatchSolve.apply((groups, rtol, max_iter, info), *dens, *loads)
    else:
        Us = [torch_solve.femsolve(d, F, g[2], g[3], g[4], g[0], g[1], rtol=rtol,
                                   max_iter=max_iter) for g, d, F in zip(groups, dens, loads)]
    out: list[Tensor | None] = [None] * len(instances)
    for (key, idxs), U, ncases in zip(items, Us, cases):
        start = 0
        for i, nc in zip(idxs, ncases):
            out[i] = U[start:start + nc]
            start += nc
    return out


def _coarsening(nelx: int, nely: int, max_coarse: int | None = None) -> list[tuple[int, int]]:
    """The coarsening factors `build_hierarchy` takes from an nelx x nely grid."""
    limit = torch_mg.MAX_COARSE_ELEMENTS if max_coarse is None else max_coarse
    out = []
    while nelx * nely > limit:
        kx, ky = torch_mg.coarsening_factors(nelx, nely)
        if (kx, ky) == (1, 1):
            break
        out.append((kx, ky))
        nelx, nely = nelx // kx, nely // ky
    return out

"""ORACLE -- test infrastructure only (imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg; never by the product).

A torch-CPU fp64 restatement of the reference's multigrid-CG algorithm, used for
(1) the exact iteration counts and V-cycle outputs of the reference algorithm on
small meshes, and (2) the CPU baseline ("port") timed beside the GPU path.  It
follows, function by function:

- matvec / diagonal / projection / operator: sttopt/torch_fem.py:62-159
  (gather, `@ KE`, density scale, `index_add_` scatter);
- pcg with row retirement and COMPACTION_RATIO: sttopt/torch_fem.py:201-343;
- coarsening, transfers, child interpolation matrices, Galerkin einsums, dense
  Cholesky of the coarsest level, V-cycle: sttopt/torch_mg.py:53-476;
- the adjoint solve with the alpha*U warm start and the element gradient:
  sttopt/torch_solve.py:117-152.

Pinned in tests/test_oracle.py: against the assembled direct oracle
(fem_direct), the Galerkin identity coarse == R A P, <Pc, f> == <c, Rf>, the
symmetry / positivity of the V-cycle, and Jacobi-PCG.
"""

from __future__ import annotations

import torch

from . import fem_direct

COMPACTION_RATIO = 0.75


class NotConverged(RuntimeError):
    def __init__(self, residuals, n_iter, rtol):
        self.residuals, self.n_iter, self.rtol = residuals, n_iter, rtol
        super().__init__(f"not converged to {rtol} in {n_iter} iterations")


def edof_tensor(nelx, nely, d=2):
    return torch.tensor(fem_direct.element_dofs(nelx, nely, d), dtype=torch.int64)


def matvec(U, density, edof, KE, ndof):
    Ue = U[..., edof]
    scaled = (Ue @ KE) * density[..., :, None]
    bs = scaled.shape[:-2]
    out = scaled.new_zeros(*bs, ndof)
    out.index_add_(-1, edof.reshape(-1), scaled.reshape(*bs, -1))
    return out


def diagonal(density, edof, KE, ndof):
    contrib = density[..., :, None] * torch.diagonal(KE)
    bs = contrib.shape[:-2]
    out = contrib.new_zeros(*bs, ndof)
    out.index_add_(-1, edof.reshape(-1), contrib.reshape(*bs, -1))
    return out


def lumped(reaction, edof, ndof):
    k = edof.shape[-1]
    contrib = reaction[..., :, None].expand(*reaction.shape, k) / 4
    bs = contrib.shape[:-2]
    out = contrib.new_zeros(*bs, ndof)
    return out.index_add(-1, edof.reshape(-1), contrib.reshape(*bs, -1))


def operator(v, density, edof, KE, ndof, mask, rdiag=None):
    v = v * mask
    Kv = matvec(v, density, edof, KE, ndof)
    if rdiag is not None:
        Kv = Kv + rdiag * v
    return Kv * mask


def safe_div(num, den):
    ok = den != 0
    return torch.where(ok, num / torch.where(ok, den, torch.ones_like(den)), torch.zeros_like(den))


def pcg(apply_A, b, apply_M, *, rtol=1e-8, max_iter=10000, x0=None, info=None):
    """Batched PCG with retirement (the reference's schedule)."""
    bnorm = b.norm(dim=-1)
    bsafe = torch.where(bnorm > 0, bnorm, torch.ones_like(bnorm))
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    r = b - apply_A(x)
    rel = r.norm(dim=-1) / bsafe
    conv = rel <= rtol
    remaining = int(torch.count_nonzero(~conv))
    nrows = b.shape[0] if b.ndim == 2 else 1
    row_iters = 0
    if remaining == 0:
        if info is not None:
            info["row_iters"] = 0
        return x, 0
    retire_ok = b.ndim == 2 and hasattr(apply_A, "select") and hasattr(apply_M, "select")
    if retire_ok:
        x_out = torch.zeros_like(b)
        rel_out = torch.zeros_like(rel)
        active = torch.arange(b.shape[0])
        last = b.shape[0]

    def compact(vs):
        nonlocal apply_A, apply_M, active, last, bsafe, rel, nrows
        done = torch.nonzero(conv).flatten()
        keep = torch.nonzero(~conv).flatten()
        x_out[active[done]] = x[done]
        rel_out[active[done]] = rel[done]
        active = active[keep]
        last = nrows = keep.numel()
        apply_A, apply_M = apply_A.select(keep), apply_M.select(keep)
        bsafe, rel = bsafe[keep], rel[keep]
        return [t[keep] for t in vs]

    if retire_ok and remaining <= COMPACTION_RATIO * last:
        x, r = compact([x, r])
    z = apply_M(r)
    p = z.clone()
    rz = (r * z).sum(-1)
    for it in range(1, max_iter + 1):
        row_iters += nrows
        Ap = apply_A(p)
        alpha = safe_div(rz, (p * Ap).sum(-1))[..., None]
        x = x + alpha * p
        r = r - alpha * Ap
        rel = r.norm(dim=-1) / bsafe
        conv = rel <= rtol
        remaining = int(torch.count_nonzero(~conv))
        if remaining == 0:
            break
        z = apply_M(r)
        rzn = (r * z).sum(-1)
        p = z + safe_div(rzn, rz)[..., None] * p
        rz = rzn
        if retire_ok and remaining <= COMPACTION_RATIO * last:
            x, r, p, rz = compact([x, r, p, rz])
    else:
        if retire_ok:
            rel_out[active] = rel
            rel = rel_out
        raise NotConverged(rel, max_iter, rtol)
    if info is not None:
        info["row_iters"] = row_iters
    if retire_ok:
        x_out[active] = x
        return x_out, it
    return x, it


def coarsening_factors(nelx, nely):
    f = lambda n: 2 if n % 2 == 0 else (3 if n % 3 == 0 else 1)  # noqa: E731
    return f(nelx), f(nely)


def interp_axis(c, k):
    if k == 1:
        return c
    w = torch.arange(k, dtype=c.dtype) / k
    body = c[..., :-1, None] * (1 - w) + c[..., 1:, None] * w
    return torch.cat([body.reshape(*c.shape[:-1], -1), c[..., -1:]], dim=-1)


def restrict_axis(f, k):
    if k == 1:
        return f
    w = torch.arange(k, dtype=f.dtype) / k
    body = f[..., :-1].reshape(*f.shape[:-1], -1, k)
    out = torch.nn.functional.pad((body * (1 - w)).sum(-1), (0, 1))
    out = out + torch.nn.functional.pad((body * w).sum(-1), (1, 0))
    out[..., -1] += f[..., -1]
    return out


def on_node_grid(v, nnx, nny, kx, ky, op):
    g = v.reshape(*v.shape[:-1], nny, nnx, v.shape[-1] // (nny * nnx))
    g = op(g.movedim(-2, -1), kx).movedim(-1, -2)
    g = op(g.movedim(-3, -1), ky).movedim(-1, -3)
    return g.reshape(*v.shape[:-1], -1)


def child_matrices(kx, ky, d=2):
    ec = torch.tensor(fem_direct.element_dofs(1, 1, d)[0])
    ef = torch.tensor(fem_direct.element_dofs(kx, ky, d))
    basis = torch.zeros(4 * d, 4 * d, dtype=torch.float64)
    basis[torch.arange(4 * d), ec] = 1.0
    fine = on_node_grid(basis, 2, 2, kx, ky, interp_axis)
    return fine[:, ef].permute(1, 2, 0).contiguous()


def child_elements(nelx, nely, kx, ky):
    e = torch.arange(nelx * nely).reshape(nely, nelx)
    return e.reshape(nely // ky, ky, nelx // kx, kx).permute(0, 2, 1, 3).reshape(-1, ky * kx)


class Level:
    def __init__(self, nelx, nely, ndof, mask, diag, edof, KE=None, density=None, keff=None,
                 rdiag=None):
        self.nelx, self.nely, self.ndof = nelx, nely, ndof
        self.mask, self.diag, self.edof = mask, diag, edof
        self.KE, self.density, self.keff, self.rdiag = KE, density, keff, rdiag
        self.kx = self.ky = 1
        self.chol = None

    def __call__(self, v):
        if self.keff is None:
            return operator(v, self.density, self.edof, self.KE, self.ndof, self.mask, self.rdiag)
        Ue = (v * self.mask)[..., self.edof]
        y = torch.einsum("...ei,...eij->...ej", Ue, self.keff)
        bs = y.shape[:-2]
        out = y.new_zeros(*bs, self.ndof)
        out.index_add_(-1, self.edof.reshape(-1), y.reshape(*bs, -1))
        return out * self.mask

    def select(self, rows):
        sel = lambda t, core: t if (t is None or t.ndim == core) else t[rows]  # noqa: E731
        lv = Level(self.nelx, self.nely, self.ndof, self.mask, sel(self.diag, 1), self.edof,
                   self.KE, sel(self.density, 1), sel(self.keff, 3), sel(self.rdiag, 1))
        lv.kx, lv.ky, lv.chol = self.kx, self.ky, sel(self.chol, 2)
        return lv


def dense_cholesky(keff, edof, ndof, mask):
    idx = (edof[:, :, None] * ndof + edof[:, None, :]).reshape(-1)
    bs = keff.shape[:-3]
    dense = keff.new_zeros(*bs, ndof * ndof)
    dense.index_add_(-1, idx, keff.reshape(*bs, -1))
    dense = dense.reshape(*bs, ndof, ndof)
    dense = (dense + dense.transpose(-1, -2)) / 2
    free = mask[:, None] & mask[None, :]
    dense = torch.where(free, dense, dense.new_zeros(()))
    dense = dense + torch.diag((~mask).to(dense.dtype))
    return torch.linalg.cholesky(dense)


def build_hierarchy(density, edof, KE, mask, nelx, nely, *, reaction=None, max_coarse_elements=700):
    ndof = mask.shape[-1]
    d = KE.shape[-1] // 4
    rdiag = None
    if reaction is not None:
        rdiag = lumped(reaction, edof, ndof)
        ME = torch.eye(KE.shape[-1], dtype=KE.dtype) / 4
    diag = diagonal(density, edof, KE, ndof)
    if rdiag is not None:
        diag = diag + rdiag
    diag = torch.where(mask, diag, torch.ones((), dtype=diag.dtype))
    levels = [Level(nelx, nely, ndof, mask, diag, edof, KE=KE, density=density, rdiag=rdiag)]
    keff = None
    while True:
        top = levels[-1]
        if top.nelx * top.nely <= max_coarse_elements:
            break
        kx, ky = coarsening_factors(top.nelx, top.nely)
        if (kx, ky) == (1, 1):
            break
        S = child_matrices(kx, ky, d)
        ch = child_elements(top.nelx, top.nely, kx, ky)
        if keff is None:
            G = torch.einsum("mpi,pq,mqj->mij", S, KE, S)
            keff = torch.einsum("...em,mij->...eij", density[..., ch], G)
            if reaction is not None:
                GM = torch.einsum("mpi,pq,mqj->mij", S, ME, S)
                keff = keff + torch.einsum("...em,mij->...eij", reaction[..., ch], GM)
        else:
            keff = torch.einsum("mpi,...empq,mqj->...eij", S, keff[..., ch, :, :], S)
        top.kx, top.ky = kx, ky
        cx, cy = top.nelx // kx, top.nely // ky
        cedof = edof_tensor(cx, cy, d)
        cndof = d * (cx + 1) * (cy + 1)
        cmask = top.mask.reshape(top.nely + 1, top.nelx + 1, d)[::ky, ::kx].reshape(-1)
        dg = torch.diagonal(keff, dim1=-2, dim2=-1)
        bs = dg.shape[:-2]
        cdiag = dg.new_zeros(*bs, cndof)
        cdiag.index_add_(-1, cedof.reshape(-1), dg.reshape(*bs, -1))
        cdiag = torch.where(cmask, cdiag, cdiag.new_ones(()))
        levels.append(Level(cx, cy, cndof, cmask, cdiag, cedof, keff=keff))
    last = levels[-1]
    if keff is None:
        keff = density[..., :, None, None] * KE
        if reaction is not None:
            keff = keff + reaction[..., :, None, None] * ME
    last.chol = dense_cholesky(keff, last.edof, last.ndof, last.mask)
    return levels


class VCycle:
    def __init__(self, levels, omega=0.6, n_smooth=2, gamma=1):
        self.levels, self.omega, self.n_smooth, self.gamma = levels, omega, n_smooth, gamma

    def select(self, rows):
        return VCycle([lv.select(rows) for lv in self.levels], self.omega, self.n_smooth, self.gamma)

    def _smooth(self, lv, x, b):
        for _ in range(self.n_smooth):
            x = x + self.omega * (b - lv(x)) / lv.diag
        return x * lv.mask

    def _cycle(self, i, b):
        lv = self.levels[i]
        if i == len(self.levels) - 1:
            x = torch.linalg.solve_triangular(lv.chol, b.unsqueeze(-1), upper=False)
            x = torch.linalg.solve_triangular(lv.chol.transpose(-1, -2), x, upper=True).squeeze(-1)
            return x * lv.mask
        x = self._smooth(lv, torch.zeros_like(b), b)
        r = b - lv(x)
        co = self.levels[i + 1]
        rc = on_node_grid(r, lv.nelx + 1, lv.nely + 1, lv.kx, lv.ky, restrict_axis) * co.mask
        ec = self._cycle(i + 1, rc)
        for _ in range(self.gamma - 1):
            ec = ec + self._cycle(i + 1, rc - co(ec))
        ef = on_node_grid(ec * co.mask, co.nelx + 1, co.nely + 1, lv.kx, lv.ky, interp_axis)
        return self._smooth(lv, x + ef * lv.mask, b)

    def __call__(self, r):
        return self._cycle(0, r * self.levels[0].mask)


def mg_solve(F, density, edof, KE, mask, nelx, nely, *, rtol=1e-8, max_iter=500, x0=None,
             reaction=None, max_coarse_elements=700, info=None, omega=0.6, n_smooth=2, gamma=1):
    """Hierarchy + MG-PCG (torch_mg.solve / FemSolve.forward); returns `(U, n_iter, levels)`."""
    levels = build_hierarchy(density, edof, KE, mask, nelx, nely, reaction=reaction,
                             max_coarse_elements=max_coarse_elements)
    M = VCycle(levels, omega, n_smooth, gamma)
    U, n = pcg(levels[0], F * mask, M, rtol=rtol, max_iter=max_iter, x0=x0, info=info)
    return U * mask, n, levels


def adjoint(levels, U, F, g, edof, KE, *, rtol=1e-8, max_iter=500, omega=0.6, n_smooth=2, gamma=1):
    """FemSolve.backward: lambda from alpha U, dL/ddensity = -(lam_e KE) . U_e."""
    mask = levels[0].mask
    alpha = safe_div((U * g).sum(-1), (U * F).sum(-1))
    x0 = alpha[..., None] * U
    lam, n = pcg(levels[0], g * mask, VCycle(levels, omega, n_smooth, gamma), rtol=rtol,
                 max_iter=max_iter, x0=x0)
    lam = lam * mask
    grad = -torch.sum((lam[..., edof] @ KE) * U[..., edof], dim=-1)
    return lam, grad, n

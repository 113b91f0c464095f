"""The Method of Moving Asymptotes on the device (drop-in for sttopt/mma.py):
`mmasub` builds the convex separable subproblem around the current point and
`subsolv` solves it with Svanberg's primal-dual interior-point Newton method.

Both run in the HIP kernels of `csrc/stt_mma.hip`.  `sttopt_mma_prepare` forms the
asymptotes, the move-limited bounds and the subproblem data P, Q, b.  The primal-dual
state stays on the device; this module drives the barrier and Newton loops of the
reference (mma.py:346-557) through `sttopt_mma_subsolv` and reads back two numbers per
line-search candidate (the residual's norm and maximum).  The (m+1) x (m+1) Newton
system is factored on the device, not the host.
"""

from __future__ import annotations

import ctypes
import warnings
from dataclasses import dataclass

import torch
from torch import Tensor

from . import _lib

_MAX_HALVINGS = 50


def trust_region_params(move_limit: Tensor, xrange: Tensor, *, asyclamp_min_ratio: float = 0.02,
                        asyclamp_max_ratio: float = 20.0) -> dict[str, Tensor]:
    """Asymptote sizing for a trust region of half-width `move_limit`
    (mma.py:22-58)."""
    asyinit = move_limit / xrange
    return {
        "asyinit": asyinit,
        "asyclamp_min": asyclamp_min_ratio * asyinit,
        "asyclamp_max": asyclamp_max_ratio * asyinit,
        "move": asyinit,
    }


@dataclass(frozen=True)
class History:
    """The last two design points and the asymptotes (mma.py:61-76)."""

    xold1: Tensor
    xold2: Tensor
    low: Tensor
    upp: Tensor

    @classmethod
    def initial(cls, xold: Tensor) -> "History":
        zeros = torch.zeros_like(xold)
        return cls(xold1=xold, xold2=xold.clone(), low=zeros, upp=zeros.clone())


def unit_box_step(history: History, iteration: int, xval: Tensor, move_limit: Tensor,
                  f0val: float, df0dx: Tensor, fval: Tensor, dfdx: Tensor, *, a0: float, c: float,
                  raa0: float, stats: dict | None = None) -> tuple[Tensor, Tensor, History]:
    """One MMA iteration on variables in [0, 1] with a trust region of half-width
    `move_limit` and every constraint relaxed at cost `c` (mma.py:79-126)."""
    m = len(fval)
    xmin = torch.zeros_like(xval)
    xmax = torch.ones_like(xval)
    zeros = torch.zeros(m, device=xval.device, dtype=xval.dtype)
    xmma, _, _, lam, *_, low, upp = mmasub(
        m, len(xval), iteration, xval, xmin, xmax, history.xold1, history.xold2, f0val, df0dx,
        fval, dfdx, history.low, history.upp, a0, zeros, torch.full_like(zeros, c),
        zeros.clone(), raa0=raa0, stats=stats, **trust_region_params(move_limit, xmax - xmin),
    )
    return xmma, lam, History(xold1=xval, xold2=history.xold1, low=low, upp=upp)


def floor_share(history: History, move_limit: Tensor, mask: Tensor) -> float:
    """Share of the variables in `mask` whose asymptotes sit at their floor
    (mma.py:129-147)."""
    floor = trust_region_params(move_limit, torch.ones_like(move_limit))["asyclamp_min"]
    at_floor = (history.upp - history.low) / 2 <= floor * (1 + 1e-6)
    return float(at_floor[mask].double().mean()) if bool(mask.any()) else 0.0


def _f64(t, n: int | None, dev) -> Tensor:
    """A contiguous fp64 device tensor of t (a tensor or a number)."""
    if not isinstance(t, Tensor):
        t = torch.full((1,) if n is None else (n,), float(t), dtype=torch.float64, device=dev)
    return t.detach().to(device=dev, dtype=torch.float64).contiguous()


def _vec_or_scalar(v, dev) -> tuple[Tensor, int]:
    """(device fp64 tensor, element stride): stride 0 for a number or a 0-d tensor."""
    if isinstance(v, Tensor) and v.ndim > 0:
        return v.detach().to(device=dev, dtype=torch.float64).contiguous(), 1
    return torch.full((1,), float(v), dtype=torch.float64, device=dev), 0


def mmasub(m: int, n: int, iteration: int, xval: Tensor, xmin: Tensor, xmax: Tensor,
           xold1: Tensor, xold2: Tensor, f0val: float, df0dx: Tensor, fval: Tensor,
           dfdx: Tensor, low: Tensor, upp: Tensor, a0: float, a: Tensor, c: Tensor, d: Tensor,
           asyinit=0.5, asyclamp_min=0.01, asyclamp_max=10.0, move=0.5, raa0: float = 1e-5,
           albefa: float = 0.1, asyincr: float = 1.2, asydecr: float = 0.7,
           stats: dict | None = None):
    """One MMA iteration: the moving asymptotes and the subproblem's solution
    (mma.py:150-276).

    :return: `(xmma, ymma, zmma, lam, xsi, eta, mu, zet, s, low, upp)`
    """
    dev = _lib.require_device(xval, xmin, xmax, df0dx, dfdx)
    out_dtype = xval.dtype
    f = lambda t: _f64(t, n, dev)  # noqa: E731
    xv, xmn, xmx, x1, x2, g0 = f(xval), f(xmin), f(xmax), f(xold1), f(xold2), f(df0dx)
    G = dfdx.detach().to(device=dev, dtype=torch.float64).reshape(m, n).contiguous()
    fv = _f64(fval, m, dev).reshape(m)
    lo = f(low).clone()
    up = f(upp).clone()
    (ai, sai), (amn, smn), (amx, smx), (mv, smv) = (_vec_or_scalar(v, dev) for v in
                                                    (asyinit, asyclamp_min, asyclamp_max, move))
    alfa, beta, p0, q0 = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(4))
    P = torch.empty(m, n, dtype=torch.float64, device=dev)
    Q = torch.empty_like(P)
    b = torch.empty(m, dtype=torch.float64, device=dev)
    lib = _lib.load()
    work = torch.empty(lib.sttopt_mma_prepare_work_count(m, n), dtype=torch.float64, device=dev)
    _lib.check(lib.sttopt_mma_prepare(
        m, n, int(iteration), xv.data_ptr(), xmn.data_ptr(), xmx.data_ptr(), x1.data_ptr(),
        x2.data_ptr(), g0.data_ptr(), fv.data_ptr(), G.data_ptr(), lo.data_ptr(), up.data_ptr(),
        ai.data_ptr(), sai, amn.data_ptr(), smn, amx.data_ptr(), smx, mv.data_ptr(), smv,
        float(raa0), float(albefa), float(asyincr), float(asydecr), alfa.data_ptr(),
        beta.data_ptr(), p0.data_ptr(), q0.data_ptr(), P.data_ptr(), Q.data_ptr(), b.data_ptr(),
        work.data_ptr(), _lib.stream_handle(dev)), "sttopt_mma_prepare")
    xmma, ymma, zmma, lam, xsi, eta, mu, zet, s = subsolv(
        m, n, 1e-7, lo, up, alfa, beta, p0, q0, P, Q, a0, a, b, c, d, stats)
    cast = lambda t: t.to(out_dtype)  # noqa: E731
    return (cast(xmma), cast(ymma), cast(zmma), cast(lam), cast(xsi), cast(eta), cast(mu),
            cast(zet), cast(s), cast(lo), cast(up))


def subsolv(m: int, n: int, epsimin: float, low: Tensor, upp: Tensor, alfa: Tensor,
            beta: Tensor, p0: Tensor, q0: Tensor, P: Tensor, Q: Tensor, a0: float, a: Tensor,
            b: Tensor, c: Tensor, d: Tensor, stats: dict | None = None):
    """The MMA subproblem by the primal-dual interior-point Newton method: barrier
    `epsi` from 1 down by 0.1 until `epsimin`, damped Newton steps with a backtracking
    line search on the residual norm (mma.py:346-557).

    :return: `(xmma, ymma, zmma, lam, xsi, eta, mu, zet, s)`
    """
    dev = _lib.require_device(low, upp, alfa, beta, p0, q0, P, Q)
    out_dtype = low.dtype
    f = lambda t, k: _f64(t, k, dev).reshape(k)  # noqa: E731
    low_, upp_, alfa_, beta_, p0_, q0_ = (f(t, n) for t in (low, upp, alfa, beta, p0, q0))
    P_ = P.detach().to(device=dev, dtype=torch.float64).reshape(m, n).contiguous()
    Q_ = Q.detach().to(device=dev, dtype=torch.float64).reshape(m, n).contiguous()
    a_, b_, c_, d_ = (f(t, m) for t in (a, b, c, d))
    lib = _lib.load()
    work = torch.empty(lib.sttopt_mma_work_count(m, n), dtype=torch.float64, device=dev)
    offs = (ctypes.c_int64 * 10)()
    _lib.check(lib.sttopt_mma_state_offsets(m, n, offs), "sttopt_mma_state_offsets")
    stream = _lib.stream_handle(dev)
    args = (m, n, low_.data_ptr(), upp_.data_ptr(), alfa_.data_ptr(), beta_.data_ptr(),
            p0_.data_ptr(), q0_.data_ptr(), P_.data_ptr(), Q_.data_ptr(), float(a0),
            a_.data_ptr(), b_.data_ptr(), c_.data_ptr(), d_.data_ptr(), work.data_ptr())
    out = work[offs[9]:offs[9] + 2]

    def op(code: int, epsi: float = 0.0, jj: int = -1) -> None:
        _lib.check(lib.sttopt_mma_subsolv(*args, code, float(epsi), int(jj), stream),
                   "sttopt_mma_subsolv")

    def residual(epsi: float, jj: int) -> tuple[float, float]:
        op(1, epsi, jj)
        norm, mx = out.tolist()
        return norm, mx

    op(0)
    epsi = 1.0
    while epsi > epsimin:
        residunorm, residumax = residual(epsi, -1)
        ittt = 0
        while residumax > 0.9 * epsi and ittt < 200:
            ittt += 1
            op(2, epsi)
            for jj in range(_MAX_HALVINGS):
                resinew, resimax = residual(epsi, jj)
                # `not >` accepts a NaN norm, as the reference does
                if not resinew > residunorm:
                    break
            op(3, epsi, jj)
            residunorm, residumax = resinew, resimax
        if stats is not None:
            stats["subsolv_newton_steps"] = stats.get("subsolv_newton_steps", 0) + ittt
        if ittt > 198:
            if stats is not None:
                stats["subsolv_cap_hits"] = stats.get("subsolv_cap_hits", 0) + 1
            warnings.warn(
                f"subsolv: inner Newton loop hit the iteration cap (ittt={ittt}) at epsi={epsi} "
                "without reaching the 0.9*epsi residual target", RuntimeWarning)
        epsi = 0.1 * epsi

    def seg(i: int, k: int) -> Tensor:
        return work[offs[i]:offs[i] + k].clone().to(out_dtype)

    x, xsi, eta = seg(0, n), seg(1, n), seg(2, n)
    y, lam, mu, s = seg(3, m), seg(4, m), seg(5, m), seg(6, m)
    z, zet = seg(7, 1).reshape(()), seg(8, 1).reshape(())
    return x, y, z, lam, xsi, eta, mu, zet, s

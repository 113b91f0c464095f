"""Density and continuity filters and the Heaviside projection of the density field
(drop-in for sttopt/filters.py).  The filter matrices are built on the host with SciPy,
as the callers expect them; applying the density filter and the projection -- forward
and backward -- runs in the HIP kernels of `csrc/stt_filter.hip`: the density filter
matrix-free as the radius-`ceil(rmin) - 1` stencil it is (`sttopt_density_filter`), the
projection elementwise (`sttopt_heaviside`).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import torch
from torch import Tensor

from . import _lib
from . import torch_util


def _neighbor_offsets(radius: float) -> list[tuple[int, int]]:
    """(di, dj) over the square of half-width `ceil(radius) - 1` (filters.py:22-25)."""
    r = int(np.ceil(radius)) - 1
    return [(di, dj) for di in range(-r, r + 1) for dj in range(-r, r + 1)]


def _element_grid(nelx: int, nely: int) -> tuple[np.ndarray, np.ndarray]:
    """(i, j) = (column, row) of every element in linear-index order (filters.py:28-35)."""
    e = np.arange(nelx * nely)
    return e % nelx, e // nelx


def density_filter(nelx: int, nely: int, rmin: float) -> tuple[sp.csr_matrix, np.ndarray]:
    """`H[e1, e2] = max(0, rmin - dist(e1, e2))` for elements within the neighbourhood
    (including e1 == e2), and the row sums `Hs` (filters.py:38-64).  The returned matrix
    is tagged with its stencil, so that `torch_util.symmetric_csr_to_tensor` and
    `apply_density_filter` apply it matrix-free."""
    i, j = _element_grid(nelx, nely)
    n = nelx * nely
    rows, cols, vals = [], [], []
    for di, dj in _neighbor_offsets(rmin):
        w = max(0.0, rmin - math.hypot(di, dj))
        if w == 0.0:
            continue
        i2, j2 = i + di, j + dj
        ok = (i2 >= 0) & (i2 < nelx) & (j2 >= 0) & (j2 < nely)
        rows.append(np.flatnonzero(ok))
        cols.append((j2 * nelx + i2)[ok])
        vals.append(np.full(int(ok.sum()), w))
    H = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    H._sttopt_stencil = (nelx, nely, float(rmin))
    Hs = np.asarray(H.sum(axis=1)).ravel()
    return H, Hs


class _FilterApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, H, Hs, x):
        ctx.H = H
        ctx.save_for_backward(Hs)
        return torch_util.sparse_apply(H, x, Hs, 1)

    @staticmethod
    def backward(ctx, g):
        (Hs,) = ctx.saved_tensors
        # (H x / Hs)' = H^T (g / Hs) = H (g / Hs)
        return None, None, torch_util.sparse_apply(ctx.H, g, Hs, 2)


def apply_density_filter(field: Tensor, H, Hs: Tensor) -> Tensor:
    """`H @ field / Hs`, shape-preserving (filters.py:67-76)."""
    return _FilterApply.apply(H, Hs, field.flatten()).reshape(field.shape)


def continuity_filter(nelx: int, nely: int, lrmin: float) -> sp.csr_matrix:
    """`L = I - D^-1 N`: each element minus the `max(0, lrmin - dist)`-weighted average of
    its neighbours, the element itself excluded (filters.py:79-106).

    :raises ValueError: if some element has no neighbour within `lrmin`
    """
    H, _ = density_filter(nelx, nely, lrmin)
    N = H.tolil()
    N.setdiag(0.0)
    N = N.tocsr()
    N.eliminate_zeros()
    row_sum = np.asarray(N.sum(axis=1)).ravel()
    if not np.all(row_sum > 0):
        raise ValueError(
            f"the continuity average of lrmin={lrmin:g} elements reaches no neighbor of some element"
        )
    n = nelx * nely
    return (sp.eye(n, format="csr") - sp.diags(1.0 / row_sum) @ N).tocsr()


def _heaviside(x: Tensor, beta: float, eta: float, derivative: bool,
               g: Tensor | None = None) -> Tensor:
    dev = _lib.require_device(x, g)
    xd = x.detach().to(torch.float64).contiguous()
    gd = None if g is None else g.detach().to(torch.float64).expand(x.shape).contiguous()
    out = torch.empty_like(xd)
    _lib.check(_lib.load().sttopt_heaviside(xd.numel(), float(beta), float(eta), int(derivative),
                                            xd.data_ptr(), _lib.ptr(gd), out.data_ptr(),
                                            _lib.stream_handle(dev)), "sttopt_heaviside")
    return out.to(x.dtype)


class _Heaviside(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, beta, eta):
        ctx.save_for_backward(x)
        ctx.beta, ctx.eta = beta, eta
        return _heaviside(x, beta, eta, False)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        if math.isinf(ctx.beta):
            return torch.zeros_like(x), None, None  # the step: zero almost everywhere
        return _heaviside(x, ctx.beta, ctx.eta, True, g), None, None


def heaviside_projection(xTilde: Tensor, beta: float, eta: float) -> Tensor:
    """`(tanh(beta eta) + tanh(beta (x - eta))) / (tanh(beta eta) + tanh(beta (1 - eta)))`;
    `beta = inf` is the step with 1/2 at `eta` (filters.py:108-125)."""
    return _Heaviside.apply(xTilde, beta, eta)


def heaviside_projection_derivative(xTilde: Tensor, beta: float, eta: float) -> Tensor:
    """d/dxTilde of `heaviside_projection` (filters.py:128-131)."""
    return _heaviside(xTilde, beta, eta, True)

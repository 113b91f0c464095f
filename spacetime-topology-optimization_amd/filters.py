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


def filter_stencil(rmin: float) -> np.ndarray:
    """The cone weights `max(0, rmin - |(di, dj)|)` on the square of half-width
    `ceil(rmin) - 1` around an element, indexed `[dj + r, di + r]` (row offset, column
    offset): the stencil `density_filter_r_kernel` applies on the device
    (filters.py:22-64)."""
    r = max(int(math.ceil(rmin)) - 1, 0)
    off = np.arange(-r, r + 1, dtype=np.float64)
    return np.clip(rmin - np.hypot(off[:, None], off[None, :]), 0.0, None)


def _stencil_matrix(nelx: int, nely: int, W: np.ndarray) -> sp.csr_matrix:
    """The element-grid operator of stencil W: the sum over its nonzero offsets of
    `W[dj, di]` times the shift to element (row + dj, column + di), a Kronecker product
    of two one-dimensional shifts (no wrap-around at the grid edges)."""
    r = W.shape[0] // 2
    n = nelx * nely
    out = sp.csr_matrix((n, n))
    for a, b in zip(*np.nonzero(W)):
        dj, di = int(a) - r, int(b) - r
        out = out + W[a, b] * sp.kron(sp.eye(nely, k=dj), sp.eye(nelx, k=di), format="csr")
    return out.tocsr()


def density_filter(nelx: int, nely: int, rmin: float) -> tuple[sp.csr_matrix, np.ndarray]:
    """`H[e1, e2] = max(0, rmin - dist(e1, e2))` for elements within the neighbourhood
    (including e1 == e2), and the row sums `Hs` (filters.py:38-64).  The returned matrix
    is tagged with its stencil, so that `torch_util.symmetric_csr_to_tensor` and
    `apply_density_filter` apply it matrix-free."""
    H = _stencil_matrix(nelx, nely, filter_stencil(rmin))
    H.eliminate_zeros()
    H.sort_indices()
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
    W = filter_stencil(lrmin)
    W[W.shape[0] // 2, W.shape[1] // 2] = 0.0  # the neighbours only
    N = _stencil_matrix(nelx, nely, W)
    N.eliminate_zeros()
    neighbour_weight = np.asarray(N.sum(axis=1)).ravel()
    if not np.all(neighbour_weight > 0):
        raise ValueError(
            f"the continuity average of lrmin={lrmin:g} elements reaches no neighbor of some element"
        )
    average = sp.diags(1.0 / neighbour_weight) @ N
    return (sp.identity(nelx * nely, format="csr") - average).tocsr()


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

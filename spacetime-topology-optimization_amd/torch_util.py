"""Boundary conversions and the sparse-matrix helpers of the density-filter path
(drop-in for sttopt/torch_util.py:19-188): NumPy to torch, SciPy CSR to
`torch.sparse_csr_tensor`, the symmetric tag, and `symmetric_matmul`, whose forward
and backward (the same product, by symmetry) run in the HIP kernels of
`csrc/stt_filter.hip`.  A matrix built by
`filters.density_filter` is recognised and applied matrix-free as the translation-
invariant stencil it is (`sttopt_density_filter`); any other symmetric CSR matrix
goes through the CSR product (`sttopt_csr_matvec`).
"""

from __future__ import annotations

import warnings
from typing import NewType

import numpy as np
import scipy.sparse as sp
import torch
from torch import Tensor

from . import _lib


def to_tensor(array, device, dtype: torch.dtype | None = None) -> Tensor:
    """`array` (a NumPy array or a tensor) as a tensor on `device`, in `dtype` when given,
    else in its own dtype (torch_util.py:19-37)."""
    if not isinstance(array, Tensor):
        array = torch.as_tensor(np.asarray(array))
    return array.to(device=device, dtype=dtype)


def to_tensors(arrays: dict, device, dtype: torch.dtype | None = None) -> dict[str, Tensor]:
    """`to_tensor` of every value of `arrays`, same keys (torch_util.py:40-59)."""
    return {key: to_tensor(value, device, dtype) for key, value in arrays.items()}


SymmetricCsr = NewType("SymmetricCsr", Tensor)
"""A `torch.sparse_csr_tensor` known to equal its own transpose (torch_util.py:127-133)."""


def csr_to_tensor(matrix, device, dtype: torch.dtype) -> Tensor:
    """SciPy sparse matrix -> `torch.sparse_csr_tensor` on `device` (sorted,
    duplicate-free columns per row; torch_util.py:62-90)."""
    csr = sp.csr_matrix(matrix)
    csr.sum_duplicates()
    csr.sort_indices()
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", "Sparse CSR tensor support is in beta state")
        t = torch.sparse_csr_tensor(
            torch.as_tensor(csr.indptr, dtype=torch.int64),
            torch.as_tensor(csr.indices, dtype=torch.int64),
            torch.as_tensor(csr.data, dtype=dtype),
            size=csr.shape,
        )
    return t.to(device)


def csr_to_scipy(x: Tensor) -> sp.csr_matrix:
    """`torch.sparse_csr_tensor` -> SciPy `csr_matrix` on the host (torch_util.py:110-124)."""
    return sp.csr_matrix(
        (x.values().detach().cpu().numpy(), x.col_indices().detach().cpu().numpy(),
         x.crow_indices().detach().cpu().numpy()),
        shape=tuple(x.shape),
    )


def to_numpy(x) -> np.ndarray:
    """A dense tensor as a host array; arrays pass through (torch_util.py:93-107)."""
    if isinstance(x, Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def symmetric_csr_to_tensor(matrix, device, dtype: torch.dtype) -> SymmetricCsr:
    """`csr_to_tensor` for a matrix that must equal its transpose, checked once
    (torch_util.py:136-154).

    :raises ValueError: if `matrix` is not symmetric
    """
    csr = sp.csr_matrix(matrix)
    if (csr != csr.T).nnz != 0:
        raise ValueError("matrix is not symmetric")
    t = csr_to_tensor(csr, device, dtype)
    spec = getattr(matrix, "_sttopt_stencil", None)
    if spec is not None:
        t._sttopt_stencil = spec  # a density-filter matrix: applied as its stencil
    return SymmetricCsr(t)


def stencil_of(matrix: Tensor) -> tuple[int, int, float] | None:
    """(nelx, nely, rmin) when `matrix` is a density-filter matrix, else None."""
    return getattr(matrix, "_sttopt_stencil", None)


def _as_rows(dense: Tensor, n: int) -> tuple[Tensor, int, tuple[int, ...]]:
    """`dense` ((n,) or (n, k)) as k contiguous fp64 rows of n, and the trailing shape."""
    if dense.shape[0] != n:
        raise ValueError(f"operand has {dense.shape[0]} rows, the matrix {n}")
    tail = tuple(dense.shape[1:])
    rows = dense.detach().to(torch.float64).reshape(n, -1).t().contiguous()
    return rows, rows.shape[0], tail


def _from_rows(rows: Tensor, tail: tuple[int, ...], dtype: torch.dtype) -> Tensor:
    return rows.t().reshape(rows.shape[1], *tail).to(dtype)


def sparse_apply(matrix: Tensor, dense: Tensor, scale: Tensor | None = None,
                 scale_mode: int = 0) -> Tensor:
    """`matrix @ dense` on the device with an optional per-entry scale (0: none, 1: divide
    the result, 2: divide the operand first), no autograd."""
    dev = _lib.require_device(matrix, dense, scale)
    n = matrix.shape[0]
    rows, k, tail = _as_rows(dense, n)
    out = torch.empty_like(rows)
    sc = None if scale is None else scale.detach().to(torch.float64).reshape(-1).contiguous()
    spec = stencil_of(matrix)
    lib = _lib.load()
    if spec is not None:
        nelx, nely, rmin = spec
        _lib.check(lib.sttopt_density_filter(nelx, nely, k, float(rmin), rows.data_ptr(), n,
                                             _lib.ptr(sc), scale_mode, out.data_ptr(), n,
                                             _lib.stream_handle(dev)), "sttopt_density_filter")
    else:
        vals = matrix.values().detach().to(torch.float64).contiguous()
        _lib.check(lib.sttopt_csr_matvec(
            n, matrix.crow_indices().contiguous().data_ptr(),
            matrix.col_indices().contiguous().data_ptr(), vals.data_ptr(), k, rows.data_ptr(), n,
            _lib.ptr(sc), scale_mode, out.data_ptr(), n, _lib.stream_handle(dev)),
            "sttopt_csr_matvec")
    return _from_rows(out, tail, dense.dtype)


class _SymmetricMatmul(torch.autograd.Function):
    @staticmethod
    def forward(ctx, matrix, dense):
        ctx.matrix = matrix
        return sparse_apply(matrix, dense)

    @staticmethod
    def backward(ctx, grad):
        # symmetric: the adjoint is the same product
        return None, sparse_apply(ctx.matrix, grad)


def symmetric_matmul(matrix: SymmetricCsr, dense: Tensor) -> Tensor:
    """`matrix @ dense` for a symmetric sparse matrix, differentiably, with the backward
    `matrix @ grad` (torch_util.py:175-188).

    :param matrix: `(n, n)` from `symmetric_csr_to_tensor`
    :param dense: `(n,)` or `(n, k)`
    """
    return _SymmetricMatmul.apply(matrix, dense)

"""Host-side bookkeeping shared by the operator modules: the structured grid behind
an element dof map, packed free-dof masks, and batch broadcasting."""

from __future__ import annotations

import torch
from torch.utils.weak import WeakTensorKeyDictionary

from . import _lib
from . import fem as _fem

# Results keyed by tensor identity and version: a caller's edofMat / mask tensor is
# validated (and a mask packed) once, not on every call.
_grid_cache = WeakTensorKeyDictionary()
_mask_cache = WeakTensorKeyDictionary()


def dofs_per_node(KE: torch.Tensor) -> int:
    k = KE.shape[-1]
    if KE.shape != (k, k) or k not in (4, 8):
        raise ValueError(f"KE must be 4x4 (scalar) or 8x8 (plane elasticity), got {tuple(KE.shape)}")
    return k // 4


def grid_of(edofMat: torch.Tensor, d: int, ndof: int | None = None) -> tuple[int, int]:
    """`(nelx, nely)` of the structured grid `edofMat` describes.

    The HIP kernels compute element/node indices from the grid instead of reading the
    map, so a map that is not the structured C-order map of `fem.element_dof_map` is
    rejected (`ValueError`) rather than silently misread.
    """
    key = (edofMat._version, d, ndof)
    hit = _grid_cache.get(edofMat)
    if hit is not None and hit[0] == key:
        return hit[1]
    nel, k = edofMat.shape
    if k != 4 * d:
        raise ValueError(f"edofMat has {k} columns, expected {4 * d} for {d} dof(s) per node")
    first = edofMat[0].detach().to("cpu").tolist()
    nelx = first[0] // d - 1  # element 0's bottom-left corner is node (1, 0)
    if nelx <= 0 or nel % nelx:
        raise ValueError("edofMat is not a structured Q4 grid dof map")
    nely = nel // nelx
    expected = torch.as_tensor(
        _fem.element_dof_map(nelx, nely, d), dtype=edofMat.dtype, device=edofMat.device
    )
    if not torch.equal(edofMat, expected):
        raise ValueError(
            "edofMat is not the structured C-order dof map of an nelx x nely grid "
            "(sttopt.fem.element_dof_map); the HIP operator only supports that layout"
        )
    if ndof is not None and ndof != d * (nelx + 1) * (nely + 1):
        raise ValueError(f"ndof={ndof} does not match a {nelx}x{nely} grid with d={d}")
    _grid_cache[edofMat] = (key, (nelx, nely))
    return nelx, nely


def mask_bits(mask: torch.Tensor) -> torch.Tensor:
    """The free-dof bool mask packed to int32 words on its device (cached)."""
    hit = _mask_cache.get(mask)
    if hit is not None and hit[0] == mask._version:
        return hit[1]
    if mask.dtype != torch.bool or mask.ndim != 1:
        raise ValueError("mask must be a 1-D bool tensor (True at free dofs)")
    dev = _lib.require_device(mask)
    n = mask.numel()
    bits = torch.empty(max((n + 31) // 32, 1), dtype=torch.int32, device=dev)
    src = mask.contiguous().view(torch.uint8)
    _lib.check(
        _lib.load().sttopt_pack_mask(src.data_ptr(), n, bits.data_ptr(), _lib.stream_handle(dev)),
        "sttopt_pack_mask",
    )
    _mask_cache[mask] = (mask._version, bits)
    return bits


def rows_of(t: torch.Tensor, core: int, batch: tuple[int, ...], dtype=None) -> tuple[torch.Tensor, int]:
    """`t` (`(*b, core)`, b broadcastable to `batch`) as a `(rows, core)` view and its
    row stride: 0 when `t` has no batch dims (shared by every row)."""
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if t.data_ptr() % 16:
        t = t.clone()  # the kernels load node pairs as 2-vectors: keep 16-byte alignment
    if t.ndim == 1:
        return t.contiguous(), 0
    if tuple(t.shape[:-1]) != batch:
        t = t.expand(*batch, core)
    t = t.reshape(-1, core)
    if not t.is_contiguous():
        t = t.contiguous()
    if t.stride(0) == 0:
        return t[0].contiguous(), 0
    return t, t.stride(0)


def batch_shape(*shapes: tuple[int, ...]) -> tuple[int, ...]:
    return tuple(torch.broadcast_shapes(*shapes))


def numel(shape: tuple[int, ...]) -> int:
    n = 1
    for s in shape:
        n *= s
    return n

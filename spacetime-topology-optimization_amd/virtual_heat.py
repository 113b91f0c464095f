"""The scalar (one dof per node) diffusion-reaction solve behind the virtual-heat
time fields: `-div(chi grad u) + reaction u = F`, `u = g` on the fixed nodes
(drop-in for the solve of sttopt/virtual_heat.py:162-288)."""

from __future__ import annotations

from dataclasses import dataclass

import torch
from torch import Tensor

from . import fem as _fem
from . import torch_solve


def element_means(u: Tensor, edof: Tensor) -> Tensor:
    """Mean of each element's four nodal values (virtual_heat.py:162-166)."""
    return u[..., edof].mean(dim=-1)


@dataclass(frozen=True)
class ScalarMesh:
    """A scalar Q4 mesh: element dof map and the diffusion element matrix."""

    nelx: int
    nely: int
    edof: Tensor
    KE: Tensor

    @classmethod
    def build(cls, nelx: int, nely: int, device, dtype=torch.float64) -> "ScalarMesh":
        return cls(
            nelx=nelx,
            nely=nely,
            edof=torch.as_tensor(_fem.element_dof_map(nelx, nely, 1), dtype=torch.int64, device=device),
            KE=torch.as_tensor(_fem.diffusion_KE(), dtype=dtype, device=device),
        )

    @property
    def ndof(self) -> int:
        return (self.nelx + 1) * (self.nely + 1)

    def solve(self, chi: Tensor, g: Tensor, fixed: Tensor, rtol: float, *,
              F: Tensor | None = None, reaction: Tensor | None = None,
              x0: Tensor | None = None, **solve_kwargs) -> Tensor:
        """`u` with `u = g` on `fixed` (virtual_heat.py:253-288)."""
        mask = ~fixed
        return torch_solve.lifted_femsolve(
            chi, torch.zeros_like(g) if F is None else F, g, self.edof, self.KE, mask,
            self.nelx, self.nely, reaction=reaction, rtol=rtol,
            x0=None if x0 is None else (x0 * mask).detach(), **solve_kwargs,
        )

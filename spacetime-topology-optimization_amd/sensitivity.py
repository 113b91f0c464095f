"""Jacobian rows of k scalar outputs w.r.t. a list of leaves (drop-in for
sttopt/sensitivity.py:15-52): one `torch.autograd.grad` per output, an exactly-zero
block for a leaf an output does not reach."""

from __future__ import annotations

from typing import Sequence

import torch
from torch import Tensor


def jacobian_rows(outputs: Tensor, leaves: Sequence[Tensor]) -> tuple[Tensor, ...]:
    """One `(k, leaf.numel())` block per leaf, in `leaves`' order."""
    leaves = tuple(leaves)
    per_output = [
        torch.autograd.grad(out, leaves, retain_graph=True, allow_unused=True)
        for out in outputs
    ]
    blocks = []
    for j, leaf in enumerate(leaves):
        rows = []
        for grads in per_output:
            g = grads[j]
            rows.append(
                torch.zeros(leaf.numel(), dtype=leaf.dtype, device=leaf.device)
                if g is None
                else g.reshape(-1)
            )
        blocks.append(torch.stack(rows))
    return tuple(blocks)

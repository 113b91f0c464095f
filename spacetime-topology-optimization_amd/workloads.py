"""Seeded synthetic problems of the solve path (mesh, loads, densities, time fields).

Every field is generated in fp64 on the host from `numpy.random.default_rng(seed)` and
then moved to the device, so the same seed gives the same inputs everywhere.

- `mbb`:        plane-stress half MBB beam, grey density clip(0.5 + 0.1 N, 1e-3, 1).
- `space_time`: bottom edge clamped, staged self-weight (+ optional top-right point
  load on the whole-structure row); density clip(mean + std N, 1e-3, 1); time field =
  white noise, Gaussian-smoothed (sigma in elements, scipy "reflect" boundary) and
  min-max scaled to [0, 1]; stage times t_k = k / n_stage.
- `heat`:       scalar steady heat with conductivity 1e-3 + rho^3, T = 0 on the bottom
  nodes, lumped unit volumetric source.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.ndimage
import torch

from . import fem as _fem

EMIN, EMAX, PENAL, NU = 1e-9, 1.0, 3.0, 0.3


def grey_density(nelx: int, nely: int, mean: float, std: float, rng) -> np.ndarray:
    return np.clip(mean + std * rng.standard_normal((nely, nelx)), 1e-3, 1.0)


def smooth_time_field(nelx: int, nely: int, sigma: float, rng) -> np.ndarray:
    t = scipy.ndimage.gaussian_filter(rng.standard_normal((nely, nelx)), sigma, mode="reflect")
    return (t - t.min()) / (t.max() - t.min())


def print_order_field(nelx: int, nely: int, sigma: float, amplitude: float, rng) -> np.ndarray:
    """A bottom-up print order: element height fraction plus `amplitude` times a
    Gaussian-smoothed (sigma elements) unit-variance noise, min-max scaled to [0, 1].
    Every stage's active set then grows from the clamped base."""
    height = (nely - 0.5 - np.arange(nely))[:, None] / nely * np.ones((1, nelx))
    n = scipy.ndimage.gaussian_filter(rng.standard_normal((nely, nelx)), sigma, mode="reflect")
    n /= n.std() + 1e-30
    t = height + amplitude * n
    return (t - t.min()) / (t.max() - t.min())


@dataclass
class Problem:
    nelx: int
    nely: int
    KE: torch.Tensor
    edofMat: torch.Tensor
    F: torch.Tensor
    freedofs: torch.Tensor
    mask: torch.Tensor
    xPhys: torch.Tensor
    ndof: int
    tPhys: torch.Tensor | None = None
    stage_times: list = field(default_factory=list)
    beta: float = 16.0
    C: object = None
    max_coarse_elements: int = 700
    name: str = ""

    @property
    def nel(self) -> int:
        return self.nelx * self.nely


def _problem(nelx, nely, case, x, device, **kw) -> Problem:
    F, free = _fem.load_case(case, nelx, nely)
    ndof = F.size
    mask = torch.zeros(ndof, dtype=torch.bool)
    mask[torch.as_tensor(free)] = True
    return Problem(
        nelx=nelx,
        nely=nely,
        KE=torch.tensor(_fem.plane_stress_KE(NU), dtype=torch.float64, device=device),
        edofMat=torch.tensor(_fem.element_dof_map(nelx, nely), dtype=torch.int64, device=device),
        F=torch.tensor(F, dtype=torch.float64, device=device),
        freedofs=torch.tensor(free, dtype=torch.int64, device=device),
        mask=mask.to(device),
        xPhys=torch.tensor(x, dtype=torch.float64, device=device),
        ndof=ndof,
        **kw,
    )


def mbb(nelx: int = 120, nely: int = 40, seed: int = 0, device="cpu",
        max_coarse_elements: int = 200) -> Problem:
    """Plain method case: MBB half beam, one fp64 MG-PCG solve."""
    rng = np.random.default_rng(seed)
    x = grey_density(nelx, nely, 0.5, 0.1, rng)
    return _problem(nelx, nely, "mbb", x, device, max_coarse_elements=max_coarse_elements,
                    name=f"mbb_{nelx}x{nely}")


def space_time(nelx: int, nely: int, n_stage: int, *, beta: float = 16.0, seed: int = 0,
               mean: float = 0.4, std: float = 0.05, sigma: float = 4.0,
               point_load: bool = True, time_field: str = "noise", amplitude: float = 0.05,
               device="cpu") -> Problem:
    """Space-time self-weight case: `1 + n_stage` rows (whole structure + stages).

    :param time_field: "noise" (smoothed white noise, sigma elements) or "print"
        (`print_order_field`: bottom-up ramp + amplitude x smoothed noise)
    """
    rng = np.random.default_rng(seed)
    x = grey_density(nelx, nely, mean, std, rng)
    if time_field == "noise":
        t = smooth_time_field(nelx, nely, sigma, rng)
    elif time_field == "print":
        t = print_order_field(nelx, nely, sigma, amplitude, rng)
    else:
        raise ValueError(f"unknown time field {time_field!r}")
    case = "bottom_clamped_tr" if point_load else "bottom_clamped"
    p = _problem(nelx, nely, case, x, device, beta=beta,
                 stage_times=[k / n_stage for k in range(1, n_stage + 1)],
                 name=f"space_time_{nelx}x{nely}_nt{n_stage}")
    p.tPhys = torch.tensor(t, dtype=torch.float64, device=device)
    p.C = 1.0 / (nelx * nely)  # fe of the unit-total-weight self-weight stencil
    return p


@dataclass
class HeatProblem:
    nelx: int
    nely: int
    KE: torch.Tensor
    edof: torch.Tensor
    chi: torch.Tensor
    F: torch.Tensor
    fixed: torch.Tensor
    rho: torch.Tensor


def heat(nelx: int = 480, nely: int = 240, seed: int = 0, device="cpu") -> HeatProblem:
    """Steady heat: chi = 1e-3 + rho^3, T = 0 on the bottom nodes, lumped unit source."""
    rng = np.random.default_rng(seed)
    rho = grey_density(nelx, nely, 0.5, 0.1, rng).ravel()
    edof = _fem.element_dof_map(nelx, nely, 1)
    nn = (nelx + 1) * (nely + 1)
    F = np.bincount(edof.ravel(), weights=np.full(edof.size, 0.25), minlength=nn)
    fixed = np.zeros(nn, dtype=bool)
    fixed[_fem.node_grid(nelx, nely)[-1, :]] = True
    return HeatProblem(
        nelx=nelx,
        nely=nely,
        KE=torch.tensor(_fem.diffusion_KE(), dtype=torch.float64, device=device),
        edof=torch.tensor(edof, dtype=torch.int64, device=device),
        chi=torch.tensor(1e-3 + rho**3, dtype=torch.float64, device=device),
        F=torch.tensor(F, dtype=torch.float64, device=device),
        fixed=torch.tensor(fixed, device=device),
        rho=torch.tensor(rho, dtype=torch.float64, device=device),
    )


UNEVEN_SIZES = [(64, 32), (128, 64), (256, 128), (512, 256)]


def smooth_density(nelx: int, nely: int, rng, volfrac: float = 0.5, sigma: float = 3.0) -> np.ndarray:
    """Gaussian-smoothed noise shifted (by bisection) to mean `volfrac` after clipping
    to [1e-3, 1]."""
    f = scipy.ndimage.gaussian_filter(rng.standard_normal((nely, nelx)), sigma, mode="reflect")
    f = (f - f.mean()) / (f.std() + 1e-30) * 0.2
    lo, hi = -1.0, 2.0
    for _ in range(60):
        mid = (lo + hi) / 2
        if np.clip(f + mid, 1e-3, 1.0).mean() < volfrac:
            lo = mid
        else:
            hi = mid
    return np.clip(f + (lo + hi) / 2, 1e-3, 1.0)


def uneven_batch(n_instances: int = 64, seed: int = 0, device="cpu", sizes=None):
    """`n_instances` cantilever instances (left edge clamped) with grid sizes drawn from
    `sizes`, each with two load cases: a uniform -y traction over the whole top edge
    and a unit point load at a seeded free node in a seeded direction.

    :return: list of dicts with nelx, nely, x (nely, nelx), loads (2, ndof), freedofs
    """
    rng = np.random.default_rng(seed)
    sizes = UNEVEN_SIZES if sizes is None else sizes
    out = []
    for _ in range(n_instances):
        nelx, nely = sizes[rng.integers(len(sizes))]
        x = smooth_density(nelx, nely, rng)
        _, free = _fem.load_case("cantilever", nelx, nely)
        nodes = _fem.node_grid(nelx, nely)
        ndof = 2 * nodes.size
        loads = np.zeros((2, ndof))
        loads[0, 2 * nodes[0, :] + 1] = -_fem.edge_load_shares(nelx + 1, nelx)
        node = int(rng.integers(nelx + 1, nodes.size))  # off the clamped first column
        while node % (nelx + 1) == 0:
            node = int(rng.integers(nelx + 1, nodes.size))
        ang = rng.uniform(0, 2 * np.pi)
        loads[1, 2 * node] = np.cos(ang)
        loads[1, 2 * node + 1] = np.sin(ang)
        out.append(dict(nelx=nelx, nely=nely, x=x, loads=loads, freedofs=free))
    return out

"""The scalar (one dof per node) Q4 mesh behind the virtual-heat time fields
(drop-in for sttopt/virtual_heat.py:51-462): the diffusion-reaction solve
`-div(chi grad u) + reaction u = F`, `u = g` on the fixed nodes; the Gauss-point
gradient fields (`gauss_gradients`, `direction_field`), the divergence load and the
Poisson distance fitted to a direction field; and the heat and Laplace time fields
built from them.  The gradient fields, the load and their adjoints run in the HIP kernels of
`csrc/stt_heat.hip` (`sttopt_gauss_gradients`, `sttopt_divergence_load`,
`sttopt_gauss_dot`, `sttopt_direction_field_backward`); the solves in the batched
MG-PCG of `torch_solve`.
"""

from __future__ import annotations

import ctypes
import enum
import warnings
from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np
import scipy.sparse as sp
import torch
from scipy.sparse.csgraph import dijkstra
from torch import Tensor

from . import _lib
from . import fem as _fem
from . import torch_solve

#: Floor on the penalized density in `chi = density**p * chi(mu)`
#: (virtual_heat.py:42-44).
DENSITY_FLOOR = 1e-6

#: Floor on `T` in the `neg_log` map, relative to the smallest `T` on the part
#: (virtual_heat.py:46-48).
NEG_LOG_FLOOR = 1e-3


class LaplaceTimeMap(str, enum.Enum):
    """How the harmonic field maps to the time field (the values of the reference's
    `run_config.LaplaceTimeMap`, run_config.py:487-491)."""

    IDENTITY = "identity"
    POISSON = "poisson"


class HeatTimeMap(str, enum.Enum):
    """How the virtual temperature `T` maps to the time field (the values of the
    reference's `run_config.HeatTimeMap`, run_config.py:479-485)."""

    ONE_MINUS = "one_minus"
    NEG_LOG = "neg_log"
    POISSON = "poisson"


def _gauss_points() -> list[tuple[float, float]]:
    """(xi, eta) of the 2x2 Gauss rule on the unit element, xi fastest."""
    lo, hi = 0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)
    return [(xi, eta) for eta in (lo, hi) for xi in (lo, hi)]


def _shape_gradients() -> np.ndarray:
    """`B[g, d, i]`: d/dx (d = 0, right) and d/dy (d = 1, up) of the bilinear shape
    function of local node i (BL, BR, TR, TL) at Gauss point g, per element length
    (virtual_heat.py:57-72)."""
    B = np.empty((4, 2, 4))
    for g, (xi, eta) in enumerate(_gauss_points()):
        # node: (d/dxi, d/deta) of N_BL = (1-xi)(1-eta), N_BR = xi(1-eta),
        # N_TR = xi eta, N_TL = (1-xi) eta
        per_node = [(-(1 - eta), -(1 - xi)), (1 - eta, -xi), (eta, xi), (-eta, 1 - xi)]
        for i, (dx, dy) in enumerate(per_node):
            B[g, 0, i] = dx
            B[g, 1, i] = dy
    return B


def domain_nodes(active: np.ndarray) -> np.ndarray:
    """The nodes of the design domain, (nely+1, nelx+1): those of at least one of its
    elements (virtual_heat.py:79-85)."""
    nely, nelx = active.shape
    out = np.zeros((nely + 1, nelx + 1), bool)
    for dr in (0, 1):
        for dc in (0, 1):
            out[dr:dr + nely, dc:dc + nelx] |= active
    return out


def plate_nodes(nelx: int, nely: int, print_base: str, active: np.ndarray | None = None) -> np.ndarray:
    """The build plate's nodes (flat bool over the nodes): the print-base edge on the
    design domain (virtual_heat.py:88-116).

    :raises ValueError: for a print base other than left_edge, bottom_edge, right_edge
    """
    edge = {"left_edge": (slice(None), 0), "bottom_edge": (-1, slice(None)),
            "right_edge": (slice(None), -1)}.get(print_base.lower())
    if edge is None:
        raise ValueError(
            "a virtual-heat time field needs a plate print base (left_edge, bottom_edge or "
            f"right_edge), got {print_base!r}"
        )
    plate = np.zeros((nely + 1, nelx + 1), bool)
    plate[edge] = True
    active = np.ones((nely, nelx), bool) if active is None else active
    return (plate & domain_nodes(active)).ravel()


def _boundary_loop(active: np.ndarray) -> np.ndarray:
    """The design domain's boundary nodes, each once, clockwise from its top-left node
    (virtual_heat.py:119-150).

    :raises ValueError: if the boundary is not one simple loop
    """
    nely, nelx = active.shape
    inside = np.zeros((nely + 2, nelx + 2), bool)
    inside[1:-1, 1:-1] = active
    adj: dict[tuple[int, int], list[tuple[int, int]]] = {}

    def link(a, b):
        adj.setdefault(a, []).append(b)
        adj.setdefault(b, []).append(a)

    # horizontal edge (r, c)-(r, c+1) lies between element rows r-1 and r
    for r in range(nely + 1):
        for c in range(nelx):
            if inside[r, c + 1] != inside[r + 1, c + 1]:
                link((r, c), (r, c + 1))
    # vertical edge (r, c)-(r+1, c) lies between element columns c-1 and c
    for r in range(nely):
        for c in range(nelx + 1):
            if inside[r + 1, c] != inside[r + 1, c + 1]:
                link((r, c), (r + 1, c))
    if not adj or any(len(v) != 2 for v in adj.values()):
        raise ValueError("the design domain's boundary must be one simple loop")
    start = min(adj)
    loop = [start]
    prev, cur = start, (start[0], start[1] + 1)
    while cur != start:
        loop.append(cur)
        a, b = adj[cur]
        prev, cur = cur, (b if a == prev else a)
    if len(loop) != len(adj):
        raise ValueError("the design domain's boundary must be one simple loop")
    return np.array([r * (nelx + 1) + c for r, c in loop], dtype=np.int64)


def wall_arc(nelx: int, nely: int, print_base: str, active: np.ndarray | None = None) -> np.ndarray:
    """The domain's boundary nodes off the plate, in boundary order from one end of the
    plate to the other; a corner shared with the plate belongs to the plate
    (virtual_heat.py:153-172)."""
    active = np.ones((nely, nelx), bool) if active is None else active
    loop = _boundary_loop(active)
    on = plate_nodes(nelx, nely, print_base, active)[loop]
    k = int(np.flatnonzero(on)[-1]) + 1
    ordered = np.concatenate([loop[k:], loop[:k]])
    return ordered[~plate_nodes(nelx, nely, print_base, active)[ordered]]


def element_means(u: Tensor, edof: Tensor) -> Tensor:
    """Mean of each element's four nodal values (virtual_heat.py:175-179)."""
    return u[..., edof].mean(dim=-1)


def diffusivity(mu: Tensor, contrast: float) -> Tensor:
    """`chi(mu) = contrast**(mu - 1/2)` (virtual_heat.py:182-191)."""
    return contrast ** (mu - 0.5)


def part_mask(xPhys: Tensor) -> Tensor:
    """`xPhys >= max(xPhys) / 2` (virtual_heat.py:194-200)."""
    return xPhys >= xPhys.max() / 2


def normalize(t: Tensor, xPhys: Tensor) -> Tensor:
    """`t` over its maximum on `part_mask`, the scale held constant
    (virtual_heat.py:203-213)."""
    return t / t[part_mask(xPhys)].max().detach()


# ------------------------------------------------ native gradient fields ----


def _rows(t: Tensor, core: int, batch: tuple[int, ...] | None = None) -> tuple[Tensor, int]:
    """`t` (`(*b, core)` with b broadcastable to `batch`) as fp64 rows on 16-byte
    aligned storage and its row stride: 0 when `t` has no batch dims (one row shared by
    all)."""
    t = t.detach().to(torch.float64)
    if batch is not None and t.ndim > 1 and tuple(t.shape[:-1]) != batch:
        t = t.expand(*batch, t.shape[-1])
    shared = batch is not None and t.ndim == 1 and len(batch) > 0
    t = t.reshape(-1, core)
    if not t.is_contiguous() or t.data_ptr() % 16:
        t = t.contiguous().clone()
    return t, (0 if shared else core)


def _bhost(B: Tensor) -> ctypes.Array:
    flat = B.detach().to("cpu", torch.float64).reshape(-1).tolist()
    if len(flat) != 32:
        raise ValueError(f"B must be 4 x 2 x 4, got {tuple(B.shape)}")
    arr = (ctypes.c_double * 32)()
    for i, v in enumerate(flat):
        arr[i] = v
    return arr


def _batch(*shapes) -> tuple[int, ...]:
    return tuple(torch.broadcast_shapes(*shapes))


def _numel(batch: tuple[int, ...]) -> int:
    return int(np.prod(batch)) if batch else 1


def _gauss(mesh: "ScalarMesh", u: Tensor, *, normalize: bool, escale: Tensor | None = None,
           scale: float = 1.0) -> Tensor:
    dev = _lib.require_device(u, escale)
    nel, nn = mesh.nelx * mesh.nely, mesh.ndof
    batch = _batch(tuple(u.shape[:-1]), () if escale is None else tuple(escale.shape[:-1]))
    rows = _numel(batch)
    u2, su = _rows(u, nn, batch)
    es, ses = (None, 0) if escale is None else _rows(escale, nel, batch)
    out = torch.empty(rows, nel * 8, dtype=torch.float64, device=dev)
    _lib.check(_lib.load().sttopt_gauss_gradients(
        mesh.nelx, mesh.nely, rows, mesh._B_host, u2.data_ptr(), su, _lib.ptr(es), ses,
        float(scale), int(normalize), out.data_ptr(), nel * 8, _lib.stream_handle(dev)),
        "sttopt_gauss_gradients")
    return out.reshape(*batch, nel, 4, 2)


def _load(mesh: "ScalarMesh", w: Tensor | None, X: Tensor, scale: float) -> Tensor:
    dev = _lib.require_device(w, X)
    nel, nn = mesh.nelx * mesh.nely, mesh.ndof
    batch = _batch(tuple(X.shape[:-3]), () if w is None else tuple(w.shape[:-1]))
    rows = _numel(batch)
    Xf = X.reshape(*X.shape[:-3], nel * 8)
    X2, sx = _rows(Xf, nel * 8, batch)
    w2, sw = (None, 0) if w is None else _rows(w, nel, batch)
    out = torch.empty(rows, nn, dtype=torch.float64, device=dev)
    _lib.check(_lib.load().sttopt_divergence_load(
        mesh.nelx, mesh.nely, rows, mesh._B_host, _lib.ptr(w2), sw, X2.data_ptr(), sx,
        float(scale), out.data_ptr(), nn, _lib.stream_handle(dev)), "sttopt_divergence_load")
    return out.reshape(*batch, nn)


def _gauss_dot(mesh: "ScalarMesh", X: Tensor, v: Tensor, scale: float) -> Tensor:
    dev = _lib.require_device(X, v)
    nel, nn = mesh.nelx * mesh.nely, mesh.ndof
    batch = _batch(tuple(X.shape[:-3]), tuple(v.shape[:-1]))
    rows = _numel(batch)
    X2, sx = _rows(X.reshape(*X.shape[:-3], nel * 8), nel * 8, batch)
    v2, sv = _rows(v, nn, batch)
    out = torch.empty(rows, nel, dtype=torch.float64, device=dev)
    _lib.check(_lib.load().sttopt_gauss_dot(
        mesh.nelx, mesh.nely, rows, mesh._B_host, X2.data_ptr(), sx, v2.data_ptr(), sv,
        float(scale), out.data_ptr(), nel, _lib.stream_handle(dev)), "sttopt_gauss_dot")
    return out.reshape(*batch, nel)


def _direction_backward(mesh: "ScalarMesh", s: Tensor, gY: Tensor) -> Tensor:
    dev = _lib.require_device(s, gY)
    nel, nn = mesh.nelx * mesh.nely, mesh.ndof
    batch = _batch(tuple(s.shape[:-1]), tuple(gY.shape[:-3]))
    rows = _numel(batch)
    s2, ss = _rows(s, nn, batch)
    g2, sg = _rows(gY.reshape(*gY.shape[:-3], nel * 8), nel * 8, batch)
    out = torch.empty(rows, nel * 8, dtype=torch.float64, device=dev)
    _lib.check(_lib.load().sttopt_direction_field_backward(
        mesh.nelx, mesh.nely, rows, mesh._B_host, s2.data_ptr(), ss, g2.data_ptr(), sg,
        out.data_ptr(), nel * 8, _lib.stream_handle(dev)), "sttopt_direction_field_backward")
    return out.reshape(*batch, nel, 4, 2)


def _sum_to(g: Tensor, shape: torch.Size) -> Tensor:
    """Reduce a broadcast gradient back to an input's shape."""
    while g.ndim > len(shape):
        g = g.sum(0)
    for i, n in enumerate(shape):
        if n == 1 and g.shape[i] != 1:
            g = g.sum(i, keepdim=True)
    return g


class _GaussGradients(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mesh, u):
        ctx.mesh = mesh
        return _gauss(mesh, u, normalize=False).to(u.dtype)

    @staticmethod
    def backward(ctx, gG):
        # d/du_i of sum_{e,g,d} gG B u_e = sum over the elements at i of gG . B[:, :, loc]
        return None, _load(ctx.mesh, None, gG, 1.0).to(gG.dtype)


class _DirectionField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mesh, s):
        ctx.mesh = mesh
        ctx.save_for_backward(s)
        return _gauss(mesh, s, normalize=True).to(s.dtype)

    @staticmethod
    def backward(ctx, gY):
        (s,) = ctx.saved_tensors
        gG = _direction_backward(ctx.mesh, s, gY)
        return None, _load(ctx.mesh, None, gG, 1.0).to(s.dtype)


class _DivergenceLoad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mesh, w, X):
        ctx.mesh = mesh
        ctx.save_for_backward(w, X)
        return _load(mesh, w, X, 0.25).to(X.dtype)

    @staticmethod
    def backward(ctx, gF):
        w, X = ctx.saved_tensors
        gw = gX = None
        if ctx.needs_input_grad[1]:
            # dF/dw_e: 1/4 X_e : G(gF)_e
            gw = _sum_to(_gauss_dot(ctx.mesh, X, gF, 0.25), w.shape).to(w.dtype)
        if ctx.needs_input_grad[2]:
            # dF/dX_e = 1/4 w_e G(gF)_e
            gX = _sum_to(_gauss(ctx.mesh, gF, normalize=False, escale=w, scale=0.25),
                         X.shape).to(X.dtype)
        return None, gw, gX


@dataclass(frozen=True)
class ScalarMesh:
    """The scalar Q4 mesh a virtual-heat time field is solved on
    (virtual_heat.py:216-242): element dof map, diffusion element matrix, shape-function
    gradients at the Gauss points, the plate's nodes, the walls off the plate and the
    design domain.  `B`, `plate`, `arc` and `active` may be left out by callers that
    only solve."""

    nelx: int
    nely: int
    edof: Tensor
    KE: Tensor
    B: Tensor | None = None
    plate: Tensor | None = None
    arc: Tensor | None = None
    active: Tensor | None = None
    _B_host: ctypes.Array | None = field(default=None, repr=False, compare=False)

    def __post_init__(self):
        B = self.B
        if B is None:
            B = torch.as_tensor(_shape_gradients(), dtype=self.KE.dtype, device=self.KE.device)
            object.__setattr__(self, "B", B)
        object.__setattr__(self, "_B_host", _bhost(B))

    @classmethod
    def build(cls, nelx: int, nely: int, print_base: str, device, dtype=torch.float64,
              active: np.ndarray | None = None) -> "ScalarMesh":
        """:param active: the design domain; the whole mesh if `None`"""
        act = np.ones((nely, nelx), bool) if active is None else active
        return cls(
            nelx=nelx,
            nely=nely,
            edof=torch.as_tensor(_fem.element_dof_map(nelx, nely, 1), dtype=torch.int64, device=device),
            KE=torch.as_tensor(_fem.diffusion_KE(), dtype=dtype, device=device),
            B=torch.as_tensor(_shape_gradients(), dtype=dtype, device=device),
            plate=torch.as_tensor(plate_nodes(nelx, nely, print_base, act), device=device),
            arc=torch.as_tensor(wall_arc(nelx, nely, print_base, act), dtype=torch.int64, device=device),
            active=torch.as_tensor(act.ravel(), device=device),
        )

    @property
    def ndof(self) -> int:
        return (self.nelx + 1) * (self.nely + 1)

    @property
    def unit_length(self) -> float:
        """`l_c`, in elements: the square root of the design area."""
        return (self.nelx * self.nely) ** 0.5

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

    def gauss_gradients(self, u: Tensor) -> Tensor:
        """`grad u` at each element's 2x2 Gauss points, `(*batch, nel, 4, 2)`
        (virtual_heat.py:290-292)."""
        return _GaussGradients.apply(self, u)

    def direction_field(self, s: Tensor) -> Tensor:
        """`grad s / |grad s|` at the Gauss points, with `finfo.tiny` under the root
        against 0 / 0 (virtual_heat.py:294-307)."""
        return _DirectionField.apply(self, s)

    def divergence_load(self, w: Tensor, X: Tensor) -> Tensor:
        """`F_i = integral of w X . grad N_i` by the 2x2 Gauss rule
        (virtual_heat.py:309-315)."""
        return _DivergenceLoad.apply(self, w, X)

    def poisson_distance(self, s: Tensor, w: Tensor, rtol: float, x0: Tensor | None = None) -> Tensor:
        """The distance fitted to the direction in which `s` increases:
        `div(w grad phi) = div(w X)`, `phi = 0` on the plate (virtual_heat.py:317-347)."""
        X = self.direction_field(s)
        return self.solve(w, torch.zeros_like(s), self.plate, rtol,
                          F=self.divergence_load(w, X), x0=x0)


class TimeFieldSolution(NamedTuple):
    """A virtual-heat time field and the nodal solutions behind it
    (virtual_heat.py:350-358)."""

    tPhys: Tensor
    primary: Tensor
    poisson: Tensor | None
    t: Tensor


def start_report(mesh: ScalarMesh, solution: TimeFieldSolution, solid: np.ndarray,
                 base: np.ndarray) -> dict:
    """The start check: the part touches the plate and every node of a solid element
    off the plate is later than it (virtual_heat.py:361-400)."""
    t = solution.t.detach().cpu().numpy()
    tPhys = solution.tPhys.detach().cpu().numpy().ravel()
    edof = mesh.edof.cpu().numpy()
    solid = np.asarray(solid).ravel()
    on_base = np.zeros(solid.size, bool)
    on_base[base] = True
    nodes = np.unique(edof[solid])
    nodes = nodes[~mesh.plate.cpu().numpy()[nodes]]
    earliest = float(t[nodes].min()) if nodes.size else None
    base_t = tPhys[solid & on_base]
    off_t = tPhys[solid & ~on_base]
    return dict(
        passed=bool(base_t.size) and (earliest is None or earliest > 0),
        solid_base_elements=int(base_t.size),
        earliest_node_t=earliest,
        max_base_t=float(base_t.max()) if base_t.size else None,
        earliest_off_base_t=float(off_t.min()) if off_t.size else None,
    )


def warn_start(start: dict) -> None:
    """Warn that a `start_report` failed (virtual_heat.py:403-407)."""
    warnings.warn(
        f"print start: {start['solid_base_elements']} solid base element(s), and the earliest "
        f"node of the part off the plate at t = {start['earliest_node_t']} (must be > 0)"
    )


def heat_time_field(config, mesh: ScalarMesh, xPhys: Tensor, mu: Tensor,
                    previous: TimeFieldSolution | None = None, conduction_penal: float = 1.0,
                    density_weighted: bool = True) -> TimeFieldSolution:
    """The time field from the drained heat equation through the material
    (virtual_heat.py:410-462).  `config` carries `chi_contrast`, `drain_beta`,
    `heat_cg_rtol`, `time_map` (a `HeatTimeMap` value) and `poisson_cg_rtol`."""
    shape = xPhys.shape
    if density_weighted:
        density = torch.clamp(xPhys.flatten(), min=DENSITY_FLOOR)
        penalized = torch.clamp(xPhys.flatten() ** conduction_penal, min=DENSITY_FLOOR)
    else:
        density = penalized = torch.ones_like(xPhys.flatten())
    chi = penalized * diffusivity(mu.flatten(), config.chi_contrast)
    alpha = config.drain_beta / mesh.unit_length**2
    T = mesh.solve(chi, mesh.plate.to(chi.dtype), mesh.plate, config.heat_cg_rtol,
                   reaction=torch.full_like(chi, alpha),
                   x0=None if previous is None else previous.primary)
    phi = None
    time_map = HeatTimeMap(config.time_map)
    if time_map == HeatTimeMap.ONE_MINUS:
        t = 1 - T
    elif time_map == HeatTimeMap.NEG_LOG:
        with torch.no_grad():
            floor = NEG_LOG_FLOOR * element_means(T, mesh.edof)[part_mask(xPhys).flatten()].min()
        t = -torch.log(torch.clamp(T, min=0) + floor)
    else:
        phi = mesh.poisson_distance(-T, density, config.poisson_cg_rtol,
                                    x0=None if previous is None else previous.poisson)
        t = phi
    t_e = element_means(t, mesh.edof).reshape(shape)
    return TimeFieldSolution(normalize(t_e, xPhys), T, phi, t)


# ------------------------------------------------------ Laplace time field ----


def _node_edges(active: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The node graph through the design domain (virtual_heat.py:465-489): every
    domain element links its four corners along its four sides (step 1) and across its
    two diagonals (step sqrt 2).  A side shared by two domain elements is one edge.

    :return: `(src, dst, step)` in both directions
    """
    nely, nelx = active.shape
    er, ec = np.nonzero(active)
    tl = er * (nelx + 1) + ec
    bl, br, tr = tl + nelx + 1, tl + nelx + 2, tl + 1
    pairs = [(bl, br), (tl, tr), (tl, bl), (tr, br), (bl, tr), (tl, br)]
    lengths = [1.0, 1.0, 1.0, 1.0, np.sqrt(2.0), np.sqrt(2.0)]
    a = np.concatenate([p[0] for p in pairs])
    b = np.concatenate([p[1] for p in pairs])
    w = np.concatenate([np.full(tl.size, length) for length in lengths])
    key = np.minimum(a, b) * ((nelx + 1) * (nely + 1)) + np.maximum(a, b)
    _, first = np.unique(key, return_index=True)
    a, b, w = a[first], b[first], w[first]
    return np.concatenate([a, b]), np.concatenate([b, a]), np.concatenate([w, w])


def _plate_distance(mesh: ScalarMesh) -> Tensor:
    """The wall arc's graph distance to the plate through the design domain, in units of
    `l_c` (virtual_heat.py:492-507): shortest paths over `_node_edges`, exact along the
    axes and diagonals, so on a rectangle with the plate on one edge it is the distance
    to that edge, and a wall behind a re-entrant corner takes the path around it."""
    active = mesh.active.detach().cpu().numpy().reshape(mesh.nely, mesh.nelx)
    src, dst, step = _node_edges(active)
    graph = sp.csr_matrix((step, (src, dst)), shape=(mesh.ndof, mesh.ndof))
    plate = np.flatnonzero(mesh.plate.detach().cpu().numpy())
    dist = dijkstra(graph, directed=True, indices=plate, min_only=True)
    arc = mesh.arc.detach().cpu().numpy()
    return torch.as_tensor(dist[arc] / mesh.unit_length, dtype=mesh.KE.dtype,
                           device=mesh.arc.device)


def distance_wall(mesh: ScalarMesh) -> Tensor:
    """The wall design of the distance to the plate through the domain, scaled to a
    largest value of 1/2 (virtual_heat.py:510-520)."""
    d = _plate_distance(mesh)
    return 0.5 * d / d.max()


def wall_filter(n_arc: int, rmin: float) -> np.ndarray:
    """The cone filter `max(0, rmin - |i - j|)` along the wall arc, `(n_arc, n_arc)`
    (virtual_heat.py:523-543).

    The arc meets the plate (t = 0) at positions -1 and n_arc, so the data is continued
    past each end by odd reflection about those corners (node j reappears at -2 - j and
    2 n_arc - j with its sign flipped): a ramp rising from the plate passes unchanged.
    The node's own weight is max(rmin, 1), and every row is divided by the full cone's
    sum.  rmin <= 1 leaves the data unfiltered."""
    own = max(float(rmin), 1.0)
    reach = int(np.ceil(rmin)) - 1  # farthest other node with a positive weight
    cone = lambda d: np.clip(rmin - np.abs(d), 0.0, None)  # noqa: E731
    i = np.arange(n_arc)[:, None]
    j = np.arange(n_arc)[None, :]
    H = np.where(i == j, own, cone(i - j))
    H = H - cone(i - (-2 - j)) - cone(i - (2 * n_arc - j))
    total = own + 2.0 * sum(cone(d) for d in range(1, reach + 1))
    return H / total


def laplace_time_field(config, mesh: ScalarMesh, xPhys: Tensor, mu: Tensor, b: Tensor,
                       previous: TimeFieldSolution | None = None,
                       density_exponent: float = 0.0) -> TimeFieldSolution:
    """The time field of `div(chi grad t) = 0` over the design domain with `t = 0` on the
    plate and `t = b` on the wall arc (virtual_heat.py:546-607); outside the domain
    `chi = 1` (the harmonic extension).  `config` carries `chi_contrast`,
    `chi_density_floor`, `chi_density_sensitivity`, `laplace_cg_rtol`, `time_map` (a
    `LaplaceTimeMap` value) and `poisson_cg_rtol`.

    :param density_exponent: `e` in `chi = max(xPhys**e, chi_density_floor) chi(mu)`;
        0 leaves `chi` independent of the density
    """
    shape = xPhys.shape
    chi = diffusivity(mu.flatten(), config.chi_contrast)
    if density_exponent != 0:
        rho = xPhys if config.chi_density_sensitivity else xPhys.detach()
        chi = chi * torch.clamp(rho.flatten() ** density_exponent, min=config.chi_density_floor)
    chi = torch.where(mesh.active, chi, torch.ones_like(chi))
    g = torch.zeros(mesh.ndof, dtype=chi.dtype, device=chi.device).index_put((mesh.arc,), b)
    fixed = mesh.plate.clone()
    fixed[mesh.arc] = True
    u = mesh.solve(chi, g, fixed, config.laplace_cg_rtol,
                   x0=None if previous is None else previous.primary)
    phi = None
    time_map = LaplaceTimeMap(config.time_map)
    if time_map == LaplaceTimeMap.IDENTITY:
        t = u
    else:
        phi = mesh.poisson_distance(u, torch.ones_like(chi), config.poisson_cg_rtol,
                                    x0=None if previous is None else previous.poisson)
        t = phi
    t_e = element_means(t, mesh.edof).reshape(shape)
    return TimeFieldSolution(normalize(t_e, xPhys), u, phi, t)

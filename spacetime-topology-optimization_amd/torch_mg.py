"""Geometric-multigrid hierarchy and V-cycle on the gfx950 HIP kernels.

Drop-in for the reference's `sttopt.torch_mg` (sttopt/torch_mg.py:50-535): same
coarsening rule, exact element-wise Galerkin coarse operators, projected transfers,
damped-Jacobi smoothing with equal pre/post sweeps and a dense coarsest solve, so
the V-cycle is the same SPD preconditioner.  What moves to the device:

- `build_hierarchy`: Galerkin products (`sttopt_galerkin`), coarse diagonals, the
  dense coarsest operator (`sttopt_dense_assemble`) and its inverse by the block
  symmetric sweep on the fp64 matrix cores (`sttopt_spd_inverse`); the reference's
  `chol` field is formed on first access.  Coarse element matrices are kept
  entry-major (`[row][entry][element]`) so the coarse apply reads them coalesced.
- `VCycle.__call__`: one native call (`sttopt_vcycle`).
- `solve` / `pcg(levels[0], b, VCycle)`: the whole MG-PCG in `sttopt_mgpcg`.

Precision.  The V-cycle runs in `VCYCLE_DTYPE` (fp32 by default): smoother, grid
transfers and coarse applies in fp32, the coarsest solve through the fp64 dense
inverse.  The outer PCG (operator, vectors, every inner product and norm) stays
fp64, and the fp32 V-cycle is paired with the flexible (Polak-Ribiere) beta, which
tolerates the preconditioner's rounding-level asymmetry.  Set `VCYCLE_DTYPE =
torch.float64` (or pass `dtype=`) for the all-fp64 cycle of the reference.
"""

from __future__ import annotations

import collections
import ctypes
import dataclasses
import functools
import json
import os
import threading
from dataclasses import dataclass, field
from typing import Callable

import numpy as np
import torch
from torch import Tensor
from torch.utils.weak import WeakTensorKeyDictionary

from . import _grid, _lib
from . import torch_fem

#: Stop coarsening once a level has at most this many elements (torch_mg.py:47-50).
MAX_COARSE_ELEMENTS = 700

#: Precision of the V-cycle preconditioner (the outer PCG is always fp64).
VCYCLE_DTYPE = torch.float32

#: PCG iterations between host reads of the active-row count.
CHECK_EVERY = 1

#: Queue the coarsest inverse on a second stream of the device, so that what the host
#: queues after it (the rest of the hierarchy's set-up and the solve's first V-cycle down
#: to the coarsest level) runs beside it; every user of the inverse waits for it first.
#: Same results either way.
CINV_SIDE_STREAM = True
_SIDE_STREAMS: dict = {}

#: Coarse levels whose elements cover at most this many level-0 elements are applied
#: in the V-cycle from the level-0 densities and fixed per-position matrices
#: (keff_E = sum_g rho_g G_pos(g), exact by linearity of the Galerkin product) instead
#: of reading an explicit (4d)^2 matrix per element: at 2x2 coarsening that is 4
#: densities instead of 64 matrix entries per level-1 element.
PARAM_MAX_CHILDREN = 4
#: Coarse levels whose elements cover at most this many level-0 elements are built
#: straight from the level-0 densities (density form of sttopt_galerkin: one thread per
#: coarse element runs over its block, so very deep levels lack parallelism); deeper
#: ones from the level above.
GALERKIN_MAX_BLOCK = 64


#: Time each node-stencil level's smoother schedule the first time a hierarchy shape is
#: seen (`autotune`) and use the fastest: two passes per level with the wide or the
#: narrow tile, or one pass per sweep.  All three give the same results.
AUTOTUNE = True
#: The schedules `autotune` picks between, by their flag values (sttopt_level.flags).
SCHEDULES = {
    "two_pass_wide": 0,
    "two_pass_narrow": 1 << _lib.TILE_VARIANT_SHIFT,
    "separate": _lib.LEVEL_SEPARATE,
}
#: Choices made so far: shape key -> {"levels": {level: schedule}, "us": {level:
#: {schedule: microseconds per V-cycle}}}.  Written as JSON to the file named by the
#: environment variable STTOPT_AUTOTUNE_FILE after each new choice, when it is set.
AUTOTUNE_CHOICES: dict = {}
#: A schedule name of `SCHEDULES` to use at every tunable level instead of timing.
SCHEDULE_OVERRIDE: str | None = None


#: Per-position Galerkin tables of the coarsening sequences built so far.  They depend
#: on the reference element matrix and the coarsening factors only, not on the
#: densities, so one hierarchy build per step reuses them.  Keyed by the KE tensor
#: object and its version counter (an in-place change invalidates); each entry keeps
#: its KE alive, so the id cannot be reused while cached.  A new table is complete on
#: the device before it is shared (its stream is synchronized once), and the device is
#: synchronized before an entry is dropped, so tables are safe to read from any stream
#: (the uneven-batch solver runs buckets on their own streams and threads).
_TABLES: collections.OrderedDict = collections.OrderedDict()
_TABLES_MAX = 32
_TABLES_LOCK = threading.Lock()


def _cached_table(KE: Tensor, key: tuple, make):
    full = (id(KE), KE._version, key)
    with _TABLES_LOCK:
        hit = _TABLES.get(full)
        if hit is not None and hit[0] is KE:
            _TABLES.move_to_end(full)
            return hit[1]
        value = make()
        if value.is_cuda:
            torch.cuda.current_stream(value.device).synchronize()
        _TABLES[full] = (KE, value)
        if len(_TABLES) > _TABLES_MAX:
            if value.is_cuda:
                torch.cuda.synchronize(value.device)
            while len(_TABLES) > _TABLES_MAX:
                _TABLES.popitem(last=False)
        return value


def coarsening_factors(nelx: int, nely: int) -> tuple[int, int]:
    """Per-dimension factor: the smallest of 2, 3 dividing it, else 1 (torch_mg.py:53-72)."""

    def one(n: int) -> int:
        for k in (2, 3):
            if n % k == 0:
                return k
        return 1

    return one(nelx), one(nely)


@functools.lru_cache(maxsize=None)
def _hat_matrix(nc: int, k: int, dtype, device) -> Tensor:
    """One-dimensional prolongation on a line of nodes, `(k nc + 1, nc + 1)`: fine node i
    lies at coarse coordinate i / k and takes the coarse nodal hat functions' values
    there, `max(0, 1 - |i / k - j|)` (linear interpolation, exact at coarse nodes)."""
    fine = torch.arange(k * nc + 1, dtype=torch.float64)[:, None] / k
    coarse = torch.arange(nc + 1, dtype=torch.float64)[None, :]
    return torch.clamp(1 - (fine - coarse).abs(), min=0).to(device=device, dtype=dtype)


def _interp_axis(c: Tensor, k: int) -> Tensor:
    """Linear interpolation by factor `k` along the last axis, `nc+1 -> k*nc+1` nodes
    (torch_mg.py:75-82): the hat-function matrix applied to that axis."""
    if k == 1:
        return c
    P = _hat_matrix(c.shape[-1] - 1, k, c.dtype, c.device)
    return c @ P.T


def _restrict_axis(f: Tensor, k: int) -> Tensor:
    """Full weighting along the last axis, the exact transpose of `_interp_axis`
    (torch_mg.py:85-95)."""
    if k == 1:
        return f
    P = _hat_matrix((f.shape[-1] - 1) // k, k, f.dtype, f.device)
    return f @ P


def _along(t: Tensor, axis: int, op, k: int) -> Tensor:
    """`op(., k)` applied along `axis` of `t` (op works on the last axis)."""
    return op(t.movedim(axis, -1), k).movedim(-1, axis)


def _on_node_grid(v: Tensor, nnx: int, nny: int, kx: int, ky: int, axis_op) -> Tensor:
    """`axis_op` along x, then along y, of a dof vector laid out as an `(nny, nnx, d)`
    node grid with `d` interleaved dofs per node (torch_mg.py:98-109)."""
    lead = v.shape[:-1]
    grid = v.reshape(*lead, nny, nnx, v.shape[-1] // (nnx * nny))
    grid = _along(_along(grid, -2, axis_op, kx), -3, axis_op, ky)
    return grid.reshape(*lead, -1)


@functools.lru_cache(maxsize=None)
def element_dof_map_tensor(nelx: int, nely: int, device, dofs_per_node: int = 2) -> Tensor:
    """The C-order element dof map (`fem.element_dof_map`) as an int64 device tensor,
    cached per grid and device (torch_mg.py:112-127): element (r, c) lists its corners
    BL, BR, TR, TL, each with its `d` interleaved dofs."""
    d = dofs_per_node
    r = torch.arange(nely, dtype=torch.int64)[:, None]
    c = torch.arange(nelx, dtype=torch.int64)[None, :]
    top_left = (r * (nelx + 1) + c).reshape(-1, 1)
    corners = top_left + torch.tensor([nelx + 1, nelx + 2, 1, 0], dtype=torch.int64)
    dofs = d * corners[:, :, None] + torch.arange(d, dtype=torch.int64)
    return dofs.reshape(nelx * nely, 4 * d).to(device)


@functools.lru_cache(maxsize=None)
def _child_elements(nelx: int, nely: int, kx: int, ky: int, device) -> Tensor:
    """`(nel_coarse, kx*ky)`: the fine elements under each coarse element, children in
    row-major order within the block (torch_mg.py:130-137).  Coarse element (R, C) holds
    fine element (R ky + a, C kx + b) as child a kx + b."""
    R = torch.arange(nely // ky)[:, None, None, None]
    C = torch.arange(nelx // kx)[None, :, None, None]
    a = torch.arange(ky)[None, None, :, None]
    b = torch.arange(kx)[None, None, None, :]
    fine = (R * ky + a) * nelx + C * kx + b
    return fine.reshape(-1, ky * kx).to(device)


@functools.lru_cache(maxsize=None)
def child_interpolation_matrices(
    kx: int, ky: int, device=None, dtype=torch.float64, dofs_per_node: int = 2
) -> Tensor:
    """`S[m]`, `(kx*ky, 4d, 4d)`: child element m's nodal values from its coarse
    element's under bilinear interpolation (torch_mg.py:140-165).

    In the coarse element's local coordinates (x right, y up, both in [0, 1]), child
    m = a kx + b (block row a from the top, column b) has its corners BL, BR, TR, TL at
    x in {b, b + 1} / kx and y in {ky - a - 1, ky - a} / ky; entry [p d + i, q d + j] is
    the coarse shape function of corner q at child corner p when i == j (each dof
    component interpolates on its own)."""
    d = dofs_per_node
    corner_x = np.array([0.0, 1.0, 1.0, 0.0])  # BL, BR, TR, TL
    corner_y = np.array([0.0, 0.0, 1.0, 1.0])

    def shape(x, y):  # N_q(x, y) of the coarse corners, q in BL, BR, TR, TL order
        return np.where(corner_x > 0, x, 1 - x) * np.where(corner_y > 0, y, 1 - y)

    S = np.zeros((kx * ky, 4 * d, 4 * d))
    for a in range(ky):
        for b in range(kx):
            for p in range(4):
                x = (b + corner_x[p]) / kx
                y = (ky - a - 1 + corner_y[p]) / ky
                S[a * kx + b, p * d:(p + 1) * d, :] = np.kron(shape(x, y), np.eye(d))
    return torch.as_tensor(S, dtype=dtype, device=device).contiguous()


def _to_soa(keff: Tensor) -> Tensor:
    """(*b, nel, k, k) -> (rows, k*k, nel) entry-major."""
    nel, k = keff.shape[-3], keff.shape[-1]
    return keff.reshape(-1, nel, k * k).transpose(1, 2).contiguous()


def _soa_view(soa: Tensor, k: int, dbatch: tuple[int, ...]) -> Tensor:
    """(rows, k*k, nel) -> (*dbatch, nel, k, k) view."""
    rows, _, nel = soa.shape
    return soa.reshape(rows, k, k, nel).permute(0, 3, 1, 2).reshape(*dbatch, nel, k, k)


def keff_matvec(U: Tensor, keff: Tensor, edofMat: Tensor, ndof: int) -> Tensor:
    """Matvec of a level with one matrix per element (torch_mg.py:179-193)."""
    dev = _lib.require_device(U, keff)
    k = keff.shape[-1]
    d = k // 4
    nelx, nely = _grid.grid_of(edofMat, d, ndof)
    soa = _to_soa(keff.to(U.dtype))
    kb = tuple(keff.shape[:-3])
    batch = _grid.batch_shape(U.shape[:-1], kb)
    rows = _grid.numel(batch)
    if soa.shape[0] not in (1, rows):
        soa = _to_soa(keff.expand(*batch, *keff.shape[-3:]).to(U.dtype))
    sk = 0 if soa.shape[0] == 1 else soa.stride(0)
    u2, su = _grid.rows_of(U, ndof, batch)
    out = torch.empty(rows, ndof, dtype=U.dtype, device=dev)
    _lib.check(
        _lib.load().sttopt_apply_keff(
            _lib.dtype_code(U.dtype), d, nelx, nely, rows, u2.data_ptr(), su,
            soa.data_ptr(), sk, None, out.data_ptr(), ndof, _lib.stream_handle(dev),
        ),
        "sttopt_apply_keff",
    )
    return out.reshape(*batch, ndof)


def _select_rows(t: Tensor | None, rows: Tensor, core_ndim: int) -> Tensor | None:
    if t is None or t.ndim == core_ndim:
        return t
    if t.ndim != core_ndim + 1:
        raise AssertionError(f"row selection needs a single batch dim, got shape {tuple(t.shape)}")
    return t[rows]


@dataclass
class _Level:
    """One grid of the hierarchy; calling it applies `A(v) = P(K P v)` on that grid
    (torch_mg.py:211-266).  Level 0 carries `density` (and `reaction_diag`); coarser
    levels carry `keff`, a `(*batch, nel, k, k)` view of the entry-major storage.  A
    level that the fp32 V-cycle applies from the level-0 densities forms that storage
    only on first access (`keff`, the fp64 cycle, `__call__`, `select`)."""

    nelx: int
    nely: int
    ndof: int
    mask: Tensor
    diag: Tensor
    edofMat: Tensor
    KE: Tensor | None = None
    density: Tensor | None = None
    _keff: Tensor | None = field(default=None, repr=False)
    reaction_diag: Tensor | None = None
    kx: int = 1
    ky: int = 1
    # coarsest level: the assembled dense operator (*batch, n, n) and its Cholesky factor
    # `chol`, both formed on first access of `chol` (the V-cycle uses `_cinv`)
    _dense: Tensor | None = field(default=None, repr=False)
    _chol: Tensor | None = field(default=None, repr=False)
    # device storage used by the kernels
    _soa: Tensor | None = field(default=None, repr=False)  # (rows, k*k, nel) fp64
    _cinv: Tensor | None = field(default=None, repr=False)  # (rows, n, n) fp64
    # the inverse's not-positive-definite word on the device, until it has been read
    _cinv_status: Tensor | None = field(default=None, repr=False)
    # the event recorded after the inverse's launches on the side stream, until a user
    # of the inverse has made its stream wait for it
    _cinv_ready: object = field(default=None, repr=False)
    _dbatch: tuple = field(default=(), repr=False)
    _copies: dict = field(default_factory=dict, repr=False)
    # per-position matrices of the level-0 elements under one element of this level
    # ((pky*pkx, k*k) fp64) and the block size (pkx, pky); None with a reaction term
    _ptab: Tensor | None = field(default=None, repr=False)
    _pblock: tuple = field(default=(1, 1), repr=False)
    # builds `_soa` on first access (levels formed lazily), else None
    _soa_make: Callable[[], Tensor] | None = field(default=None, repr=False)

    @property
    def keff(self) -> Tensor | None:
        """The element matrices, `(*batch, nel, k, k)`; None on level 0."""
        self._materialize()
        return self._keff

    @property
    def coarse(self) -> bool:
        """Whether the level is applied from element matrices (every level but 0)."""
        return self._keff is not None or self._soa_make is not None

    def _materialize(self) -> None:
        if self._soa_make is not None:
            self._soa = self._soa_make()
            self._soa_make = None
            self._keff = _soa_view(self._soa, 4 * self.d, self._dbatch)

    @property
    def d(self) -> int:
        return self.ndof // ((self.nelx + 1) * (self.nely + 1))

    @property
    def rows(self) -> int:
        return _grid.numel(self._dbatch)

    @property
    def chol(self) -> Tensor | None:
        """Lower Cholesky factor of the coarsest level's dense operator (the reference's
        `_Level.chol`, torch_mg.py:397), computed on first access."""
        if self._chol is None and self._dense is None and self._cinv is not None:
            dense = _assemble_dense(self, max(self.rows, 1))
            self._dense = dense.reshape(*self._dbatch, self.ndof, self.ndof)
        if self._chol is None and self._dense is not None:
            self._chol = torch.linalg.cholesky(self._dense)
        return self._chol

    def bits(self) -> Tensor:
        return _grid.mask_bits(self.mask)

    def cinv_wait(self) -> None:
        """Make the current stream wait for the coarsest inverse when it was queued on
        the side stream (`_dense_factor`)."""
        ready = self._cinv_ready
        if ready is not None:
            torch.cuda.current_stream(self.mask.device).wait_event(ready)
            self._cinv_ready = None

    def spd_check(self) -> None:
        """Raise the coarsest factorization's error if its inverse, queued without a wait
        (`_dense_factor`), found a matrix that is not positive definite; reads the device
        word once."""
        self.cinv_wait()
        status = self._cinv_status
        if status is None:
            return
        if int(status.item()) != 0:
            raise torch.linalg.LinAlgError(_NOT_SPD)
        self._cinv_status = None

    def __call__(self, v: Tensor) -> Tensor:
        if not self.coarse:
            return torch_fem.operator(
                v, self.density, self.edofMat, self.KE, self.ndof, self.mask, self.reaction_diag
            )
        self._materialize()
        dev = _lib.require_device(v)
        batch = _grid.batch_shape(v.shape[:-1], self._dbatch)
        rows = _grid.numel(batch)
        soa = self._soa.to(v.dtype)
        if soa.shape[0] not in (1, rows):
            raise NotImplementedError("level rows must be 1 or match the vector batch")
        sk = 0 if soa.shape[0] == 1 else soa.stride(0)
        u2, su = _grid.rows_of(v, self.ndof, batch)
        out = torch.empty(rows, self.ndof, dtype=v.dtype, device=dev)
        _lib.check(
            _lib.load().sttopt_apply_keff(
                _lib.dtype_code(v.dtype), self.d, self.nelx, self.nely, rows, u2.data_ptr(),
                su, soa.data_ptr(), sk, self.bits().data_ptr(), out.data_ptr(), self.ndof,
                _lib.stream_handle(dev),
            ),
            "sttopt_apply_keff",
        )
        return out.reshape(*batch, self.ndof)

    def select(self, rows: Tensor) -> "_Level":
        """This level restricted to batch rows `rows` (shared tensors pass through)."""
        self._materialize()
        soa = self._soa
        dbatch = self._dbatch
        if len(dbatch) == 1:
            dbatch = (int(rows.numel()),)
            if soa is not None:
                soa = soa[rows]
        self.cinv_wait()
        cinv = _select_rows(self._cinv, rows, 2) if self._cinv is not None and len(self._dbatch) == 1 else self._cinv
        k = 4 * self.d
        return dataclasses.replace(
            self,
            diag=_select_rows(self.diag, rows, 1),
            density=_select_rows(self.density, rows, 1),
            reaction_diag=_select_rows(self.reaction_diag, rows, 1),
            _keff=None if soa is None else _soa_view(soa, k, dbatch),
            _dense=_select_rows(self._dense, rows, 2),
            _chol=_select_rows(self._chol, rows, 2),
            _soa=soa,
            _cinv=cinv,
            _dbatch=dbatch,
            _copies={},
        )

    # V-cycle data in a given precision, cached per level
    def vdata(self, dtype: torch.dtype) -> dict:
        hit = self._copies.get(dtype)
        if hit is not None:
            return hit
        rows = max(self.rows, 1)
        out = {}
        dev = self.mask.device
        diag = self.diag.detach().to(torch.float64).reshape(rows, self.ndof)
        inv = torch.empty(rows, self.ndof, dtype=dtype, device=dev)
        _lib.check(
            _lib.load().sttopt_invert_diagonal(
                _lib.dtype_code(dtype), self.ndof, rows, diag.data_ptr(), self.ndof,
                self.bits().data_ptr(), inv.data_ptr(), self.ndof, _lib.stream_handle(dev),
            ),
            "sttopt_invert_diagonal",
        )
        out["invdiag"] = inv
        if not self.coarse:
            out["density"] = self.density.detach().to(dtype).reshape(rows, -1).contiguous()
            if self.reaction_diag is not None:
                out["rdiag"] = self.reaction_diag.detach().to(dtype).reshape(rows, -1).contiguous()
        elif self._ptab is not None and self._pblock[0] * self._pblock[1] <= PARAM_MAX_CHILDREN:
            out["ptable"] = self._ptab.to(dtype).contiguous()
        else:
            self._materialize()
            if dtype == torch.float64:
                out["keff"] = self._soa
            else:
                # the fp32 V-cycle applies the assembled operator as a node stencil: each
                # node-pair coupling stored once (one 20-value record per node at d = 2
                # against 36 values per element for the packed element matrices)
                d = self.d
                nn = (self.nelx + 1) * (self.nely + 1)
                nrow = self._soa.shape[0]
                node = torch.empty(nrow, nn, 4 + 4 * d * d, dtype=dtype, device=dev)
                _lib.check(
                    _lib.load().sttopt_node_stencil(
                        _lib.dtype_code(dtype), d, self.nelx, self.nely, nrow,
                        self._soa.data_ptr(), self._soa.stride(0), node.data_ptr(),
                        node.stride(0), _lib.stream_handle(dev),
                    ),
                    "sttopt_node_stencil",
                )
                out["keff"] = node
                out["keff_layout"] = _lib.KEFF_NODE
        self._copies[dtype] = out
        return out


def _assemble_dense(level: _Level, rows: int) -> Tensor:
    """The coarsest level's dense projected operator, `(rows, n, n)` fp64
    (torch_mg._dense_cholesky's assembly, torch_mg.py:269-290)."""
    dev = level.mask.device
    n = level.ndof
    dense = torch.empty(rows, n, n, dtype=torch.float64, device=dev)
    keep = None
    if level.coarse:
        level._materialize()
        args = (level._soa.data_ptr(), level._soa.stride(0), None, 0, None, 0, None)
    else:
        den = level.density.detach().to(torch.float64).reshape(rows, -1).contiguous()
        rd = level.reaction_diag
        rd = None if rd is None else rd.detach().to(torch.float64).reshape(rows, -1).contiguous()
        ke_dev = level.KE.detach().to(device=dev, dtype=torch.float64).contiguous()
        keep = (den, rd, ke_dev)
        args = (None, 0, den.data_ptr(), den.stride(0), _lib.ptr(rd),
                0 if rd is None else rd.stride(0), ke_dev.data_ptr())
    _lib.check(
        _lib.load().sttopt_dense_assemble(
            level.d, level.nelx, level.nely, rows, *args, level.bits().data_ptr(),
            dense.data_ptr(), n * n, _lib.stream_handle(dev),
        ),
        "sttopt_dense_assemble",
    )
    # the temporaries in `keep` may be freed now: the caching allocator hands their
    # memory only to later work on this stream
    del keep
    return dense


_NOT_SPD = ("linalg.cholesky: The factorization could not be completed because the input is "
            "not positive-definite (the coarsest multigrid operator).")


def _side_stream(dev: torch.device) -> torch.cuda.Stream:
    """The second stream of `dev` (default priority) that `_dense_factor` queues on."""
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    side = _SIDE_STREAMS.get(idx)
    if side is None:
        side = _SIDE_STREAMS[idx] = torch.cuda.Stream(device=torch.device("cuda", idx))
    return side


def _dense_factor(level: _Level, rows: int) -> None:
    """Inverse of the coarsest level's dense projected operator (torch_mg._dense_cholesky,
    torch_mg.py:269-290; the inverse replaces the factor's two triangular solves per
    coarse solve, torch_mg.py:448-451).  The operator is assembled into the buffer that
    is then inverted in place; `chol` assembles it again if a caller asks for it.

    The inverse is queued without a wait for its positive-definiteness check, so that the
    host goes on queueing the build and the solve while it runs: the solvers read the
    check with their final readback (`sttopt_mg.cinv_status`) and every other user of the
    inverse calls `spd_check` first.  A matrix that is not positive definite raises
    torch.linalg.LinAlgError (as the reference's `torch.linalg.cholesky`) from there.
    With `CINV_SIDE_STREAM` the inverse runs on a second stream of the device (its sweep
    steps are bound by one workgroup each and leave most of the device free): the solvers'
    V-cycles wait for its event (`sttopt_mg.cinv_ready`) before their first coarsest
    solve, and the other users through `cinv_wait`.
    """
    dev = level.mask.device
    n = level.ndof
    inv = _assemble_dense(level, rows)
    lib = _lib.load()
    work = torch.empty(lib.sttopt_spd_inverse_work_count(n, rows), dtype=torch.float64,
                       device=dev)
    status = torch.empty(1, dtype=torch.int32, device=dev)
    # the operator couples a node with its eight grid neighbours only: half-bandwidth
    # d (nnx + 1) + d - 1 in the node-major dof numbering
    band = level.d * (level.nelx + 2) + level.d - 1
    side = _side_stream(dev) if CINV_SIDE_STREAM else None
    if side is not None:
        side.wait_stream(torch.cuda.current_stream(dev))
        # allocated on the current stream, used on side: marked before anything is
        # queued there, so that no block returns to the allocator while side may use it
        for t in (inv, work, status):
            t.record_stream(side)
    stream = side.cuda_stream if side is not None else _lib.stream_handle(dev)
    _lib.check(lib.sttopt_spd_inverse_band_async(n, band, rows, inv.data_ptr(), n * n,
                                                 work.data_ptr(), status.data_ptr(), stream),
               "sttopt_spd_inverse")
    level._cinv_ready = None
    if side is not None:
        ready = torch.cuda.Event()
        ready.record(side)
        level._cinv_ready = ready
    level._cinv = inv
    level._cinv_status = status
    level._dense = None
    level._chol = None


_PTAB_DIAG_HOST: dict = {}


def _param_diagonal(den64: Tensor, ptab: Tensor, cmask: Tensor, nelx: int, nely: int,
                    kx: int, ky: int, d: int) -> Tensor:
    """Jacobi diagonal `(rows, ndof_c)` fp64 of the level whose element matrices are
    sum_g rho_g ptab[gy, gx] over the kx x ky level-0 elements under each element
    (`sttopt_param_diagonal`; 1 at fixed dofs, as `sttopt_keff_diagonal` gives it).  The
    positions' diagonals are read back once per cached table."""
    rows = den64.shape[0]
    cnelx, cnely = nelx // kx, nely // ky
    hit = _PTAB_DIAG_HOST.get(id(ptab))
    if hit is None or hit[0] is not ptab:
        diag = torch.diagonal(ptab, dim1=-2, dim2=-1).reshape(-1).to("cpu", torch.float64)
        host = (ctypes.c_double * diag.numel())(*diag.tolist())
        hit = _PTAB_DIAG_HOST[id(ptab)] = (ptab, host)  # keeps ptab, so its id stays unique
        while len(_PTAB_DIAG_HOST) > _TABLES_MAX:
            _PTAB_DIAG_HOST.pop(next(iter(_PTAB_DIAG_HOST)))
    out = torch.empty(rows, d * (cnelx + 1) * (cnely + 1), dtype=torch.float64,
                      device=den64.device)
    _lib.check(
        _lib.load().sttopt_param_diagonal(
            d, cnelx, cnely, kx, ky, rows, den64.data_ptr(), den64.stride(0), hit[1],
            _grid.mask_bits(cmask).data_ptr(), out.data_ptr(), out.stride(0),
            _lib.stream_handle(den64.device),
        ),
        "sttopt_param_diagonal",
    )
    return out


# The coarse free-dof masks of a fine mask, per coarsening: (the fine mask's version,
# the coarse mask).  Hierarchies built on one mask (every step of an optimization) share
# them, and so their packed bits (`_grid.mask_bits`, cached per mask tensor).
_COARSE_MASKS = WeakTensorKeyDictionary()


def _coarse_mask(mask: Tensor, nelx: int, nely: int, kx: int, ky: int, d: int) -> Tensor:
    """The free-dof mask of the kx x ky coarsening of a grid's mask: the nodes at every
    kx-th column and ky-th row (torch_mg.py:376)."""
    per = _COARSE_MASKS.get(mask)
    key = (nelx, nely, kx, ky, d)
    if per is not None and per[0] == mask._version and key in per[1]:
        return per[1][key]
    cmask = mask.reshape(nely + 1, nelx + 1, d)[::ky, ::kx].reshape(-1).contiguous()
    if per is None or per[0] != mask._version:
        per = _COARSE_MASKS[mask] = (mask._version, {})
    per[1][key] = cmask
    return cmask


def build_hierarchy(
    density: Tensor,
    edofMat: Tensor,
    KE: Tensor,
    mask: Tensor,
    nelx: int,
    nely: int,
    *,
    reaction: Tensor | None = None,
    max_coarse_elements: int = MAX_COARSE_ELEMENTS,
) -> list[_Level]:
    """The multigrid levels for one density field or a batch of them, finest first
    (torch_mg.py:293-400), built on the device."""
    dev = _lib.require_device(density, mask, reaction)
    ndof = mask.shape[-1]
    d = _grid.dofs_per_node(KE)
    if _grid.grid_of(edofMat, d, ndof) != (nelx, nely):
        raise ValueError(f"edofMat does not describe a {nelx}x{nely} grid")
    dbatch = tuple(density.shape[:-1])
    rows = max(_grid.numel(dbatch), 1)
    nel = nelx * nely
    den64 = density.detach().to(torch.float64).reshape(rows, nel).contiguous()
    react64 = None
    reaction_diag = None
    if reaction is not None:
        react64 = reaction.detach().to(torch.float64).expand(*dbatch, nel).reshape(rows, nel).contiguous()
        reaction_diag = torch_fem.lumped_reaction(reaction.detach(), edofMat, ndof)
    diag = torch_fem.jacobi_preconditioner_diag(
        density.detach(), edofMat, KE, ndof, mask, reaction_diag
    )
    KE64 = KE.detach().to(device=dev, dtype=torch.float64)
    levels = [
        _Level(nelx=nelx, nely=nely, ndof=ndof, mask=mask, diag=diag, edofMat=edofMat,
               KE=KE, density=density.detach(), reaction_diag=reaction_diag, _dbatch=dbatch)
    ]
    lib = _lib.load()
    k = 4 * d
    ptab = None  # (Ky, Kx, k, k) per-position matrices of the current coarsest level
    seq: tuple = ()  # the coarsening factors so far (the tables' cache key)
    while True:
        top = levels[-1]
        if top.nelx * top.nely <= max_coarse_elements:
            break
        kx, ky = coarsening_factors(top.nelx, top.nely)
        if (kx, ky) == (1, 1):
            break
        first = len(levels) == 1
        S = child_interpolation_matrices(kx, ky, dev, torch.float64, d)
        cnelx, cnely = top.nelx // kx, top.nely // ky
        # per-position matrices of the new level: level-0 element (gy, gx) of the block
        # under one coarse element contributes density * ptab[gy, gx]
        seq = seq + ((kx, ky),)
        if first:
            ptab = _cached_table(
                KE, ("ptab", seq, str(dev), d),
                lambda: torch.einsum("mpi,pq,mqj->mij", S, KE64, S).reshape(ky, kx, k, k))
        elif ptab is not None:
            def deeper(prev=ptab, S=S, kx=kx, ky=ky):
                Ky, Kx = prev.shape[:2]
                nxt = torch.empty(ky, Ky, kx, Kx, k, k, dtype=torch.float64, device=dev)
                for my in range(ky):
                    for mx in range(kx):
                        Sm = S[my * kx + mx]
                        nxt[my, :, mx] = torch.einsum("pi,yxpq,qj->yxij", Sm, prev, Sm)
                return nxt.reshape(ky * Ky, kx * Kx, k, k)

            ptab = _cached_table(KE, ("ptab", seq, str(dev), d), deeper)
        density_form = first or (react64 is None and ptab is not None
                                 and ptab.shape[0] * ptab.shape[1] <= GALERKIN_MAX_BLOCK)
        if density_form:
            # straight from the level-0 densities (any depth)
            GM = None
            if first and react64 is not None:
                GM = _cached_table(
                    KE, ("reaction", (kx, ky), str(dev), d),
                    lambda: torch.einsum("mpi,pq,mqj->mij", S,
                                         torch.eye(k, dtype=torch.float64, device=dev) / 4,
                                         S).contiguous())
            table = ptab.reshape(-1, k * k).contiguous()

            def make(table=table, GM=GM, bx=ptab.shape[1], by=ptab.shape[0],
                     n_el=cnelx * cnely) -> Tensor:
                out = torch.empty(rows, k * k, n_el, dtype=torch.float64, device=dev)
                _lib.check(
                    lib.sttopt_galerkin(
                        d, nelx, nely, bx, by, rows, den64.data_ptr(), nel, _lib.ptr(react64),
                        nel, table.data_ptr(), _lib.ptr(GM), None, 0, None, out.data_ptr(),
                        out.stride(0), _lib.stream_handle(dev),
                    ),
                    "sttopt_galerkin",
                )
                return out
        else:
            top._materialize()
            fine = top._soa

            def make(fine=fine, S=S, fnelx=top.nelx, fnely=top.nely, kx=kx, ky=ky,
                     n_el=cnelx * cnely) -> Tensor:
                out = torch.empty(rows, k * k, n_el, dtype=torch.float64, device=dev)
                _lib.check(
                    lib.sttopt_galerkin(
                        d, fnelx, fnely, kx, ky, rows, None, 0, None, 0, None, None,
                        fine.data_ptr(), fine.stride(0), S.data_ptr(), out.data_ptr(),
                        out.stride(0), _lib.stream_handle(dev),
                    ),
                    "sttopt_galerkin",
                )
                return out
        if react64 is not None:
            ptab = None  # the reaction term has no per-position form below level 1
        top.kx, top.ky = kx, ky
        cndof = d * (cnelx + 1) * (cnely + 1)
        cmask = _coarse_mask(top.mask, top.nelx, top.nely, kx, ky, d)
        # the first level of an fp32 cycle is applied from the level-0 densities
        # (PARAM_MAX_CHILDREN): its element matrices are formed only if something
        # reads them, and its diagonal comes from the per-position tables
        lazy = first and ptab is not None and kx * ky <= PARAM_MAX_CHILDREN
        if lazy:
            soa = None
            cdiag = _param_diagonal(den64, ptab, cmask, nelx, nely, kx, ky, d)
        else:
            soa = make()
            cdiag = torch.empty(rows, cndof, dtype=torch.float64, device=dev)
            _lib.check(
                lib.sttopt_keff_diagonal(
                    _lib.F64, d, cnelx, cnely, rows, soa.data_ptr(), soa.stride(0),
                    _grid.mask_bits(cmask).data_ptr(), 0, cdiag.data_ptr(), cndof,
                    _lib.stream_handle(dev),
                ),
                "sttopt_keff_diagonal",
            )
        levels.append(
            _Level(
                nelx=cnelx, nely=cnely, ndof=cndof, mask=cmask,
                diag=cdiag.reshape(*dbatch, cndof).to(density.dtype),
                edofMat=element_dof_map_tensor(cnelx, cnely, dev, d),
                _keff=None if soa is None else _soa_view(soa, k, dbatch), _soa=soa,
                _soa_make=make if lazy else None, _dbatch=dbatch,
                _ptab=None if ptab is None else ptab.reshape(-1, k * k).contiguous(),
                _pblock=(1, 1) if ptab is None else (ptab.shape[1], ptab.shape[0]),
            )
        )
    _dense_factor(levels[-1], rows)
    return levels


class VCycle:
    """Multigrid V-cycle as an SPD preconditioner `z = M(r)` (torch_mg.py:403-476).

    :param levels: `build_hierarchy` output, finest first
    :param omega: damped-Jacobi factor
    :param n_smooth: pre- and post-smoothing sweeps per level
    :param gamma: coarse visits per level (1: V-cycle, 2: W-cycle)
    :param dtype: V-cycle precision (default `VCYCLE_DTYPE`)
    :param separate_sweeps: run every smoothing sweep as its own pass instead of the
        fused passes (same results; for testing)
    """

    def __init__(self, levels, *, omega: float = 0.6, n_smooth: int = 2, gamma: int = 1,
                 dtype: torch.dtype | None = None, separate_sweeps: bool = False) -> None:
        self.levels = levels
        self.separate_sweeps = separate_sweeps
        self.omega = omega
        self.n_smooth = n_smooth
        self.gamma = gamma
        self.dtype = VCYCLE_DTYPE if dtype is None else dtype
        self._work: dict = {}

    def select(self, rows: Tensor) -> "VCycle":
        return VCycle([lv.select(rows) for lv in self.levels], omega=self.omega,
                      n_smooth=self.n_smooth, gamma=self.gamma, dtype=self.dtype,
                      separate_sweeps=self.separate_sweeps)

    # -- native descriptor --------------------------------------------------

    def _native(self, rows: int):
        """The `sttopt_mg` descriptor for `rows` batch rows, with every buffer it
        points at kept alive on this object."""
        hit = self._work.get(rows)
        if hit is not None:
            return hit
        lv0 = self.levels[0]
        hrows = max(lv0.rows, 1)
        if hrows not in (1, rows):
            raise NotImplementedError(
                f"hierarchy built for {hrows} rows cannot precondition {rows} rows"
            )
        shared = hrows == 1 and rows > 1
        dev = lv0.mask.device
        vd = self.dtype
        mg = _lib.MG()
        mg.rows, mg.d, mg.nlevels = rows, lv0.d, len(self.levels)
        mg.vdtype = _lib.dtype_code(vd)
        mg.n_smooth, mg.gamma, mg.omega = self.n_smooth, self.gamma, self.omega
        ke = _lib.ke_array(lv0.KE)
        for i in range(64):
            mg.ke[i] = ke[i]
        keep = []
        den64 = lv0.density.detach().to(torch.float64).reshape(hrows, -1).contiguous()
        keep.append(den64)
        mg.density64 = den64.data_ptr()
        mg.density64_stride = 0 if shared else den64.stride(0)
        if lv0.reaction_diag is not None:
            rd64 = lv0.reaction_diag.detach().to(torch.float64).reshape(hrows, -1).contiguous()
            keep.append(rd64)
            mg.rdiag64 = rd64.data_ptr()
            mg.rdiag64_stride = 0 if shared else rd64.stride(0)
        for l, lv in enumerate(self.levels):
            L = mg.levels[l]
            L.nelx, L.nely, L.kx, L.ky = lv.nelx, lv.nely, lv.kx, lv.ky
            if l == 0 and self.separate_sweeps:
                L.flags = _lib.SEPARATE_SWEEPS
            bits = lv.bits()
            keep.append(bits)
            L.mask = bits.data_ptr()
            data = lv.vdata(vd)
            inv = data["invdiag"]
            L.invdiag = inv.data_ptr()
            L.invdiag_stride = 0 if shared else inv.stride(0)
            if "density" in data:
                L.density = data["density"].data_ptr()
                L.density_stride = 0 if shared else data["density"].stride(0)
                if "rdiag" in data:
                    L.rdiag = data["rdiag"].data_ptr()
                    L.rdiag_stride = 0 if shared else data["rdiag"].stride(0)
            if "keff" in data:
                L.keff = data["keff"].data_ptr()
                L.keff_stride = 0 if shared else data["keff"].stride(0)
                L.keff_layout = data.get("keff_layout", _lib.KEFF_FULL)
            if "ptable" in data:
                L.ptable = data["ptable"].data_ptr()
                L.pkx, L.pky = lv._pblock
            if lv._cinv is not None:
                L.cinv = lv._cinv.data_ptr()
                L.cinv_stride = 0 if shared else lv._cinv.stride(0)
                if lv._cinv_status is not None:  # read by the solvers' final readback
                    keep.append(lv._cinv_status)
                    mg.cinv_status = lv._cinv_status.data_ptr()
                if lv._cinv_ready is not None:  # waited for before the first coarsest solve
                    keep.append(lv._cinv_ready)
                    mg.cinv_ready = lv._cinv_ready.cuda_event
            names = ["vb", "vx", "vx2", "vr"] + (["ve", "vw"] if self.gamma > 1 else [])
            for name in names:
                buf = torch.empty(rows, lv.ndof, dtype=vd, device=dev)
                keep.append(buf)
                setattr(L, name, buf.data_ptr())
        nbytes = _lib.load().sttopt_mg_scratch_bytes(ctypes.byref(mg))
        scratch = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
        keep.append(scratch)
        mg.scratch = scratch.data_ptr()
        work = torch.empty(_lib.load().sttopt_mgpcg_work_count(ctypes.byref(mg)),
                           dtype=torch.float64, device=dev)
        if SCHEDULE_OVERRIDE is not None and not self.separate_sweeps:
            for l in _tunable_levels(self):
                mg.levels[l].flags = SCHEDULES[SCHEDULE_OVERRIDE]
        elif AUTOTUNE and not self.separate_sweeps:
            for l, flags in autotune(self, mg, rows).items():
                mg.levels[l].flags = flags
        entry = (mg, keep, work)
        self._work[rows] = entry
        return entry

    def __call__(self, r: Tensor) -> Tensor:
        dev = _lib.require_device(r)
        lv0 = self.levels[0]
        self.levels[-1].spd_check()
        batch = _grid.batch_shape(r.shape[:-1], lv0._dbatch)
        rows = max(_grid.numel(batch), 1)
        mg, _, _ = self._native(rows)
        r2, rs = _grid.rows_of(r, lv0.ndof, batch, torch.float64)
        z = torch.empty(rows, lv0.ndof, dtype=torch.float64, device=dev)
        _lib.check(
            _lib.load().sttopt_vcycle(ctypes.byref(mg), r2.data_ptr(), rs, z.data_ptr(),
                                      lv0.ndof, _lib.stream_handle(dev)),
            "sttopt_vcycle",
        )
        return z.reshape(*batch, lv0.ndof).to(r.dtype)

    # -- native PCG (torch_fem.pcg fast path) ----------------------------------

    def _native_pcg(self, apply_A, b: Tensor, *, rtol, max_iter, x0, info):
        """`torch_fem.pcg(levels[0], b, self)` entirely in `sttopt_mgpcg`; None when
        `apply_A` is not this hierarchy's level 0."""
        if apply_A is not self.levels[0] or b.dtype != torch.float64:
            return None
        return native_mgpcg(self, b, rtol=rtol, max_iter=max_iter, x0=x0, info=info)


def _tunable_levels(vc: VCycle) -> list[int]:
    """Levels whose smoother schedule has a choice: node-stencil levels of the fp32 cycle
    coarsened 2 x 2 with two sweeps and one coarse visit."""
    if vc.dtype != torch.float32 or vc.n_smooth != 2 or vc.gamma != 1:
        return []
    out = []
    for l, lv in enumerate(vc.levels[:-1]):
        if l == 0 or (lv.kx, lv.ky) != (2, 2):
            continue
        data = lv.vdata(torch.float32)
        if data.get("keff_layout") == _lib.KEFF_NODE:
            out.append(l)
    return out


def autotune(vc: VCycle, mg, rows: int, trials: int = 10, repeats: int = 3) -> dict[int, int]:
    """The schedule flags of each tunable level of `vc` (`SCHEDULES`), chosen by timing
    V-cycles on a seeded residual with HIP events the first time this hierarchy shape
    and row count are seen: coarsest tunable level first, each with the coarser choices
    already in place; per level the candidates take turns, `repeats` rounds of `trials`
    cycles, and each keeps its fastest round.  Later hierarchies of the same shape reuse
    the choice."""
    levels = _tunable_levels(vc)
    if not levels:
        return {}
    key = f"d{vc.levels[0].d}_rows{rows}_" + "_".join(f"{lv.nelx}x{lv.nely}" for lv in vc.levels)
    hit = AUTOTUNE_CHOICES.get(key)
    if hit is not None:
        return {int(l): SCHEDULES[name] for l, name in hit["levels"].items()}
    lv0 = vc.levels[0]
    dev = lv0.mask.device
    g = torch.Generator(device="cpu").manual_seed(0)
    r = torch.randn(rows, lv0.ndof, generator=g, dtype=torch.float64).to(dev)
    r = r * lv0.mask.to(torch.float64)
    z = torch.empty_like(r)
    lib = _lib.load()
    stream = torch.cuda.current_stream(dev)

    def cycles(n: int) -> None:
        for _ in range(n):
            _lib.check(lib.sttopt_vcycle(ctypes.byref(mg), r.data_ptr(), lv0.ndof, z.data_ptr(),
                                         lv0.ndof, stream.cuda_stream), "sttopt_vcycle")

    def time_cycles() -> float:
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        cycles(trials)
        end.record(stream)
        end.synchronize()
        return start.elapsed_time(end) * 1e3 / trials

    cycles(3)  # warm
    chosen: dict[int, str] = {}
    timings: dict[int, dict[str, float]] = {}
    for l in reversed(levels):
        us = {name: float("inf") for name in SCHEDULES}
        for _ in range(repeats):
            for name, flags in SCHEDULES.items():
                mg.levels[l].flags = flags
                us[name] = min(us[name], time_cycles())
        best = min(us, key=us.get)
        mg.levels[l].flags = SCHEDULES[best]
        chosen[l] = best
        timings[l] = us
    AUTOTUNE_CHOICES[key] = {"levels": {str(l): n for l, n in sorted(chosen.items())},
                             "us": {str(l): t for l, t in sorted(timings.items())}}
    path = os.environ.get("STTOPT_AUTOTUNE_FILE")
    if path:
        with open(path, "w") as fh:
            json.dump(AUTOTUNE_CHOICES, fh, indent=1, sort_keys=True)
    return {l: SCHEDULES[n] for l, n in chosen.items()}


def native_mgpcg(vcycle: VCycle, b: Tensor, *, rtol: float, max_iter: int,
                 x0: Tensor | None = None, info: dict | None = None) -> tuple[Tensor, int]:
    """Batched MG-PCG of `levels[0] x = b` on the device (`sttopt_mgpcg`).

    :return: `(x, n_iter)`; x has b's shape (fixed dofs 0)
    :raises CGConvergenceError: rows above `rtol` after `max_iter`
    """
    dev = _lib.require_device(b, x0)
    lv0 = vcycle.levels[0]
    batch = _grid.batch_shape(b.shape[:-1], lv0._dbatch)
    rows = max(_grid.numel(batch), 1)
    mg, _, work = vcycle._native(rows)
    b2, bs = _grid.rows_of(b, lv0.ndof, batch, torch.float64)
    if x0 is None:
        x = torch.empty(rows, lv0.ndof, dtype=torch.float64, device=dev)  # written as 0
    else:
        x = x0.detach().to(torch.float64).expand(*batch, lv0.ndof).reshape(rows, lv0.ndof).clone()
    iters = torch.zeros(rows, dtype=torch.int32, device=dev)
    resid = torch.zeros(rows, dtype=torch.float64, device=dev)
    n_iter = ctypes.c_int32(0)
    flexible = 1 if vcycle.dtype != torch.float64 else 0
    status = _lib.load().sttopt_mgpcg(
        ctypes.byref(mg), b2.data_ptr(), bs, x.data_ptr(), lv0.ndof, float(rtol),
        int(max_iter), flexible, CHECK_EVERY, 1 if x0 is None else 0, work.data_ptr(),
        iters.data_ptr(), resid.data_ptr(), ctypes.byref(n_iter), _lib.stream_handle(dev),
    )
    if info is not None:
        per_row = iters.tolist()  # one read back
        info["row_iters"] = int(sum(per_row))
        info["row_iter_list"] = per_row
    if status == _lib.ENOTSPD:
        raise torch.linalg.LinAlgError(_NOT_SPD)
    vcycle.levels[-1]._cinv_status = None  # checked by the solve
    if status == _lib.ENOCONV:
        raise torch_fem.CGConvergenceError(resid.reshape(batch) if batch else resid[0], int(max_iter), rtol)
    _lib.check(status, "sttopt_mgpcg")
    return x.reshape(*batch, lv0.ndof), int(n_iter.value)


def solve(
    F: Tensor,
    xPhys: Tensor,
    edofMat: Tensor,
    KE: Tensor,
    Emin: float,
    Emax: float,
    penal: float,
    mask: Tensor,
    nelx: int,
    nely: int,
    *,
    rtol: float = 1e-8,
    max_iter: int = 500,
    x0: Tensor | None = None,
    omega: float = 0.6,
    n_smooth: int = 2,
    gamma: int = 1,
    max_coarse_elements: int = MAX_COARSE_ELEMENTS,
    info: dict | None = None,
) -> tuple[Tensor, int]:
    """Multigrid-preconditioned CG for `K U = F` (torch_mg.py:479-535)."""
    _lib.require_device(F, xPhys, mask, x0)
    density = torch_fem.simp_density(xPhys, Emin, Emax, penal)
    levels = build_hierarchy(density, edofMat, KE, mask, nelx, nely,
                             max_coarse_elements=max_coarse_elements)
    precond = VCycle(levels, omega=omega, n_smooth=n_smooth, gamma=gamma)
    U, n_iter = torch_fem.pcg(levels[0], torch_fem.project(F, mask), precond, rtol=rtol,
                              max_iter=max_iter, x0=x0, info=info)
    return torch_fem.project(U, mask), n_iter

"""Batched solves of independent instances of uneven size.

The reference batches rows only over one grid (one `edofMat` per call); instances of
different sizes are solved one call at a time.  Here instances are bucketed by grid
shape, and each bucket is one batched MG-PCG (`torch_solve.femsolve`, density and load
rows stacked), with the buckets issued on separate HIP streams so that the small
grids fill the GPU beside the large ones.
"""

from __future__ import annotations

from collections import OrderedDict
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import torch
from torch import Tensor

from . import _lib, torch_solve


@dataclass
class Instance:
    """One elasticity instance: grid, SIMP density `(nel,)` (requires_grad allowed),
    load cases `(n_cases, ndof)`, free-dof mask, dof map and KE on the device."""

    nelx: int
    nely: int
    density: Tensor
    loads: Tensor
    mask: Tensor
    edofMat: Tensor
    KE: Tensor


def solve_instances(instances: list[Instance], *, rtol: float = 1e-8, max_iter: int = 500,
                    info: dict | None = None, streams: bool = True) -> list[Tensor]:
    """`U` of every instance, shape `(n_cases, ndof)`, differentiable w.r.t. each
    instance's density and loads (FemSolve per bucket).

    :param info: receives `{"buckets": [(nelx, nely, rows, n_iter), ...]}`
    """
    if not instances:
        return []
    dev = _lib.require_device(*(i.density for i in instances))
    buckets: "OrderedDict[tuple, list[int]]" = OrderedDict()
    for idx, inst in enumerate(instances):
        buckets.setdefault((inst.nelx, inst.nely), []).append(idx)
    out: list[Tensor | None] = [None] * len(instances)
    main = torch.cuda.current_stream(dev)

    def run(key, idxs, stream):
        nelx, nely = key
        with torch.cuda.stream(stream):
            first = instances[idxs[0]]
            # each instance's density once per load case of its own
            ncases = [instances[i].loads.shape[0] for i in idxs]
            dens = torch.cat([instances[i].density.reshape(1, -1).expand(nc, -1)
                              for i, nc in zip(idxs, ncases)])
            loads = torch.cat([instances[i].loads for i in idxs])
            meta: dict = {}
            U = torch_solve.femsolve(dens, loads, first.edofMat, first.KE, first.mask, nelx, nely,
                                     rtol=rtol, max_iter=max_iter, info=meta)
            if stream is not main:
                U.record_stream(main)
        return U, ncases, (nelx, nely, U.shape[0], meta.get("forward_n_iter"))

    items = list(buckets.items())
    if streams and len(items) > 1:
        # one host thread and one stream per bucket: each bucket's PCG loop waits on
        # its own stream only, so the buckets overlap on the GPU (ctypes releases
        # the GIL while a native solve runs)
        pool = [torch.cuda.Stream(dev) for _ in items]
        for st in pool:
            st.wait_stream(main)
        with ThreadPoolExecutor(max_workers=len(items)) as ex:
            futs = [ex.submit(run, key, idxs, st) for (key, idxs), st in zip(items, pool)]
            results = [f.result() for f in futs]
        for st in pool:
            main.wait_stream(st)
    else:
        results = [run(key, idxs, main) for key, idxs in items]
    summary = []
    for (key, idxs), (U, ncases, meta) in zip(items, results):
        start = 0
        for i, nc in zip(idxs, ncases):
            out[i] = U[start:start + nc]
            start += nc
        summary.append(meta)
    if info is not None:
        info["buckets"] = summary
    return out

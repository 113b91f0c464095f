"""Benchmark: batched space-time stage solves + adjoint gradients per second.

Workload (one "step"): the optimizer-step hot path on a 1024 x 512 Q4 plane-stress
mesh (1,051,650 dofs): `compliance.batched_whole_and_gravity_compliance` with the
whole-structure row (top-right unit point load) and 16 smoothed-Heaviside
self-weight stages (beta 16, t_k = k/16) of a bottom-up print-order time field
(height + 0.05 x Gaussian-smoothed noise, sigma 16 elements) over a grey density
clip(0.4 + 0.05 N, 1e-3, 1) -- 17 rows solved to rel. residual 1e-8 by one batched
MG-PCG (fp64 outer iteration, fp32 V-cycle, 6 levels) -- followed by the
hand-written adjoint backward to dc/dx and dc/dt.  Inputs are seeded synthetic
fields, resident in HBM before the timed region.

value = 17 rows x steps / time  (solves with their gradients per second).

Also reported: the roofline of the dominant kernel (the fp64 K(rho)u apply of the
outer iteration) from HIP events on the solve stream, the plain fp64 one-row-at-a-
time GPU path on the same rows, and the CPU baseline (the oracle's torch-CPU fp64
restatement of the reference algorithm) on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
NELX, NELY, NSTAGE, BETA = 1024, 512, 16, 16.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nelx", type=int, default=NELX)
    ap.add_argument("--nely", type=int, default=NELY)
    ap.add_argument("--nstage", type=int, default=NSTAGE)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-batch line (64 instances of four grids)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the PCG's second half-iteration kernel by kernel instead "
                         "of replaying its captured graph")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run only a plain copy of known size and the fp64 / fp32 K(rho)u "
                         "apply (for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE), then exit")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="before the warm-up, three plain device copies of known size (the "
                         "scale of rocprofv3's FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--next-rows", action="store_true",
                    help="time the kernels around the solve (density filter, Heaviside "
                         "projection, heat-path gradient fields and load) on the same grid "
                         "and rows, print one JSON line, then exit")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path (rows shard "
              "across GPUs without communication). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import __graft_entry__

    __graft_entry__.package()
    from sttopt_amd import _lib, compliance, torch_fem, workloads

    dev = torch.device("cuda:0")
    _lib.load().sttopt_mgpcg_graphs(0 if args.no_graphs else 1)
    p = workloads.space_time(args.nelx, args.nely, args.nstage, beta=BETA, seed=0,
                             time_field="print", sigma=16.0, device=dev)
    rows = 1 + args.nstage
    if args.kernels_only:
        kernels_only(torch, _lib, p, rows, dev)
        return
    if args.next_rows:
        next_rows(torch, _lib, args.nelx, args.nely, rows, dev)
        return
    E = (workloads.EMIN, workloads.EMAX, workloads.PENAL)
    print(f"# workload {p.name}: {p.ndof} dofs x {rows} rows", file=sys.stderr, flush=True)

    def step():
        x = p.xPhys.clone().requires_grad_(True)
        t = p.tPhys.clone().requires_grad_(True)
        c, stages, U = compliance.batched_whole_and_gravity_compliance(
            x, t, p.KE, p.edofMat, *E, p.freedofs, p.F, p.ndof, p.C, p.beta, p.stage_times)
        (c + sum(stages)).backward()
        return U

    if args.calibration_copy:
        calibration_copy(torch, p, rows, dev)
    for i in range(args.warmup):
        step()
        print(f"# warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    iters = []
    lib = _lib.load()
    # HIP events around every outer-PCG K(rho)p launch of the timed steps
    _lib.check(lib.sttopt_profile_outer_apply(1), "sttopt_profile_outer_apply")
    t0 = time.perf_counter()
    for i in range(args.steps):
        U = step()
        iters.append(dict(compliance.LAST_SOLVE_INFO))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    live = [ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()]
    _lib.check(lib.sttopt_profile_outer_apply_read(*(ctypes.byref(v) for v in live)),
               "sttopt_profile_outer_apply_read")
    _lib.check(lib.sttopt_profile_outer_apply(0), "sttopt_profile_outer_apply")
    live = [v.value for v in live]
    ms_per_step = 1e3 * elapsed / args.steps
    value = rows * args.steps / elapsed
    print(f"# timed {args.steps} steps: {ms_per_step:.2f} ms/step, n_iter {iters[-1]['n_iter']}",
          file=sys.stderr, flush=True)

    # -- roofline of the dominant kernel: fp64 K(rho)p of the outer PCG -----------
    roof = apply_roofline(torch, _lib, torch_fem, p, rows, dev, live)
    print(f"# apply: {roof['achieved']:.0f} GB/s ({100 * roof['frac']:.1f}% of HBM peak)",
          file=sys.stderr, flush=True)
    # -- the whole step against its algorithmic bytes -----------------------------
    sb = step_bytes(p, iters[-1], rows)
    step_gbs = sb["total"] / (ms_per_step / 1e3) / 1e9
    roof["step"] = {"algorithmic_bytes": sb, "ms_per_step": ms_per_step,
                    "achieved": step_gbs, "frac": step_gbs / HBM_PEAK_GBS}
    print(f"# step: {sb['total'] / 1e9:.2f} GB algorithmic, {step_gbs:.0f} GB/s "
          f"({100 * step_gbs / HBM_PEAK_GBS:.1f}% of HBM peak)", file=sys.stderr, flush=True)
    resid = true_residuals(torch, p, U, E)
    print(f"# true fp64 residual per row: max {max(resid):.2e}", file=sys.stderr, flush=True)

    # -- the plain fp64 one-row-at-a-time GPU path -------------------------------
    seq = None
    if not args.no_sequential:
        seq = sequential_fp64(torch, p, rows, E)
        print(f"# fp64 one-row-at-a-time: {seq['rows_per_s']:.2f} rows/s", file=sys.stderr,
              flush=True)

    # -- the uneven batch: 64 instances of four grids x 2 load cases ----------------
    uneven_line = None
    if not args.no_uneven:
        uneven_line = uneven_throughput(torch, dev)
        print(f"# uneven batch: {uneven_line['solves_per_s']:.1f} solves+grads/s in one batched "
              f"MG-PCG, {uneven_line['per_grid_calls_solves_per_s']:.1f} with one call per grid; "
              f"forward solves alone {uneven_line['forward_solves_only']['batched_ms']:.2f} ms "
              f"batched, {uneven_line['forward_solves_only']['per_grid_calls_ms']:.2f} ms per grid",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, E)
        print(f"# cpu baseline: {cpu['value']:.4f} rows/s on {cpu['cores']} threads",
              file=sys.stderr, flush=True)

    out = {
        "metric": "batched stage solves+adjoint grads/s, 1024x512 Q4, rtol 1e-8; % HBM peak",
        "value": value,
        "unit": "solves+grads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"space-time self-weight {args.nelx}x{args.nely} Q4, 1 whole-structure "
                        f"row + {args.nstage} stages (beta {BETA}, print-order time field), "
                        "batched MG-PCG rtol 1e-8 + hand-written adjoint",
            "rows": rows,
            "ndof": p.ndof,
            "preconditioner": "fp32 V-cycle (damped Jacobi 2+2, omega 0.6), fp64 outer PCG",
            "pcg_iterations": iters[-1]["n_iter"],
            "row_iterations": iters[-1]["row_iter_list"],
            "true_relative_residual_per_row": resid,
            "levels": iters[-1]["levels"],
            "fp64_sequential_rows_per_s": None if seq is None else seq["rows_per_s"],
            "speedup_vs_fp64_sequential": None if seq is None else value / seq["rows_per_s"],
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "uneven_batch": uneven_line,
    }
    print(json.dumps(out), flush=True)


def step_bytes(p, info, rows):
    """Algorithmic HBM bytes of one bench step: the compulsory traffic of every kernel
    (each operand read once, each result written once, at its stored precision;
    per-row scalars and reduction partials left out), summed over the kernels the step
    launches.  Returns {"build", "per_row_iteration", "row_iterations", "iterations",
    "gradient", "total"}.

    The fp32 V-cycle with fp64 outer PCG on 2 dofs per node; level l has n_l dofs,
    m_l nodes and e_l elements.  Per PCG iteration of a row still iterating:
      outer apply          8 (2 n0 + e0)                      p, rho -> Ap
      pre pass + update    8 (3 n0) + 4 (4 n0 + e0)           r, Ap, z_old, 1/D, rho32 -> r', x, r32
      level 1 (densities)  two-sweep pass 4 (3 n1) + 4 e0, residual 4 (3 n1) + 4 e0,
                           two post sweeps 2 (4 (4 n1) + 4 e0)
      levels 2..L-2        (node stencils, 80 B per node) the same with 80 m_l in place
                           of 4 e0; levels above 16384 nodes take a first sweep
                           4 (3 n_l) and a plain sweep instead of the two-sweep pass
      transfers            restrict 4 (n_l + n_l+1), prolongate 4 (2 n_l + n_l+1)
      coarsest             8 n_L^2 + 4 (2 n_L)                  dense inverse, b -> x
      post pass            8 n0 + 4 (3 n0 + n1 + e0)          r, x, e, 1/D, rho32 -> z
      direction update     8 (2 n0) + 4 n0 + 8 (3 n0) / 2     p, z -> p; x, p_prev -> x
                           (the x updates of two iterations in one pass over x)
    Per solve: the first preconditioner application (the cycle without the residual
    update, 8 n0 less) and z -> p (12 n0), the cold start (24 n0); per step: the
    prologue 8 (2 e0 + n0 + R (e0 + n0)), the hierarchy (fp32 densities 12 e0, the
    Galerkin element matrices 8 * 64 e_l written for levels 2.., their node stencils
    512 e_l + 80 m_l, diagonals 8 n_l, the coarsest operator 512 e_L + 8 n_L^2 and its
    in-place inverse 16 n_L^2, per row), the compliance 8 R (n0 + e0) and the adjoint
    8 (R n0 + 4 e0)."""
    lv = info["levels"]
    nodes = [(x + 1) * (y + 1) for x, y in lv]
    n = [2 * m for m in nodes]
    e = [x * y for x, y in lv]
    L = len(lv) - 1
    e0 = e[0]

    def vcycle_below_0():
        b = 0
        for lvl in range(1, L):
            op = 4 * e0 if lvl == 1 else 80 * nodes[lvl]
            fused = lvl == 1 or nodes[lvl] <= 16384
            b += (4 * 3 * n[lvl] + op) if fused else (4 * 3 * n[lvl] + 4 * 4 * n[lvl] + op)
            b += 4 * 3 * n[lvl] + op                       # residual
            b += 2 * (4 * 4 * n[lvl] + op)                 # post sweeps
            b += 4 * (n[lvl] + n[lvl + 1])                 # restrict to lvl + 1
            b += 4 * (2 * n[lvl] + n[lvl + 1])             # prolongate from lvl + 1
        b += 4 * (n[0] + n[1])                             # restrict from level 0
        b += 8 * n[L] ** 2 + 4 * 2 * n[L]                  # coarsest
        return b

    pre = 8 * 3 * n[0] + 4 * (4 * n[0] + e0)
    post = 8 * n[0] + 4 * (3 * n[0] + n[1] + e0)
    direction = 8 * 2 * n[0] + 4 * n[0] + 8 * 3 * n[0] // 2
    per_iter = 8 * (2 * n[0] + e0) + pre + vcycle_below_0() + post + direction
    first = (pre - 8 * n[0]) + vcycle_below_0() + post + 12 * n[0] + 24 * n[0]
    build = 8 * (2 * e0 + n[0] + rows * (e0 + n[0]))          # prologue
    per_row_build = 12 * e0 + 8 * n[0]
    for lvl in range(1, L + 1):
        per_row_build += 8 * n[lvl]
        if lvl >= 2:
            per_row_build += 8 * 64 * e[lvl]
            if lvl < L:
                per_row_build += 512 * e[lvl] + 80 * nodes[lvl]
    per_row_build += 512 * e[L] + 8 * n[L] ** 2 + 16 * n[L] ** 2
    build += rows * (per_row_build + first)
    grad = 8 * rows * (n[0] + e0) + 8 * (rows * n[0] + 4 * e0)
    row_iters = int(sum(info["row_iter_list"]))
    total = build + row_iters * per_iter + grad
    return {"build": build, "per_row_iteration": per_iter, "row_iterations": row_iters,
            "iterations": int(info["n_iter"]), "gradient": grad, "total": total}


def true_residuals(torch, p, U, E):
    """Per-row relative residual |P (F_row - K_row U_row)| / |P F_row| in fp64, recomputed
    from the final U with the fp64 operator (the densities and loads rebuilt on the
    device as the fused prologue builds them)."""
    from sttopt_amd import compliance, torch_fem

    rows = U.shape[0]
    out = []
    for k in range(rows):
        if k == 0:
            dens = torch_fem.simp_density(p.xPhys, *E).flatten()
            F = p.F
        else:
            m = compliance.time_mask(p.tPhys, p.stage_times[k - 1], p.beta)
            xt = (p.xPhys * m).flatten()
            dens = torch_fem.simp_density(xt, *E)
            F = torch.zeros(p.ndof, dtype=torch.float64, device=U.device)
            F[1::2] = -(torch_sparse_gravity(torch, p) @ xt)
        Fp = torch_fem.project(F, p.mask)
        r = Fp - torch_fem.operator(U[k], dens, p.edofMat, p.KE, p.ndof, p.mask)
        out.append(float(r.norm() / Fp.norm()))
    return out


PMC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                           "stencil_tile_f64_1024x512_pmc.md")


def _apply_args(torch, _lib, p, rows, dev, dtype, code, mode=None):
    """Arguments of one sttopt_apply launch over all rows.  Default mode: the outer
    PCG's apply (input already zero at fixed dofs, output projected)."""
    from sttopt_amd import _grid

    mode = _lib.MODE_OPERATOR_PROJECTED if mode is None else mode
    g = torch.Generator(device="cpu").manual_seed(0)
    u = torch.randn(rows, p.ndof, generator=g, dtype=torch.float64).to(dev, dtype)
    u = u * p.mask.to(dtype)  # a PCG direction: zero at the fixed dofs
    dens = (torch.rand(rows, p.nel, generator=g, dtype=torch.float64) + 0.1).to(dev, dtype)
    y = torch.empty_like(u)
    bits = _grid.mask_bits(p.mask)
    stream = torch.cuda.current_stream(dev)
    args = (code, mode, 2, p.nelx, p.nely, rows, u.data_ptr(), p.ndof,
            dens.data_ptr(), p.nel, _lib.ke_array(p.KE), bits.data_ptr(), None, 0,
            y.data_ptr(), p.ndof, stream.cuda_stream)
    return args, (u, dens, y, bits), stream


def calibration_copy(torch, p, rows, dev):
    """Three plain device copies (torch `copy_`) of rows x (2 ndof + nel) fp64 values:
    what rocprofv3's FETCH_SIZE / WRITE_SIZE read for them, against the bytes they
    move, scales the counters of the other kernels to bytes."""
    n = rows * (2 * p.ndof + p.nel)
    src = torch.randn(n, device=dev, dtype=torch.float64)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"# copy: {n * 8} bytes read, {n * 8} bytes written per launch", file=sys.stderr,
          flush=True)


def kernels_only(torch, _lib, p, rows, dev):
    """The counter calibration copy, then the fp64 and fp32 K(rho)u applies of the
    outer PCG on all rows: 5 launches each, then 20 timed with HIP events."""
    lib = _lib.load()
    calibration_copy(torch, p, rows, dev)
    for dtype, code in ((torch.float64, _lib.F64), (torch.float32, _lib.F32)):
        args, keep, stream = _apply_args(torch, _lib, p, rows, dev, dtype, code)
        for _ in range(5):
            _lib.check(lib.sttopt_apply(*args), "sttopt_apply")
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(20):
            lib.sttopt_apply(*args)
        end.record(stream)
        end.synchronize()
        us = start.elapsed_time(end) * 1e3 / 20
        nbytes = rows * dtype.itemsize * (2 * p.ndof + p.nel)
        print(f"# apply {dtype} {p.nelx}x{p.nely} x {rows} rows: {us:.1f} us, "
              f"{nbytes / us / 1e3:.0f} GB/s ({nbytes} algorithmic bytes)", flush=True)
        del keep
    torch.cuda.synchronize()


def next_rows(torch, _lib, nelx, nely, rows, dev, launches=50, rmin=3.0):
    """HIP-event timings of the kernels around the solve on `rows` fields of the
    nelx x nely grid: the density filter (fp64, radius rmin, scaled by Hs), the
    Heaviside projection, and the scalar mesh's direction field and divergence load;
    and the wall clock of one MMA step (mmasub) with rows constraints over the
    2 nel space-time design variables.
    Algorithmic bytes per launch: filter 8 (x + Hs + y) per element, projection
    8 (x + y), direction field 8 per node + 64 per element, load 8 (w) + 64 (X) per
    element + 8 per node."""
    from sttopt_amd import virtual_heat

    lib = _lib.load()
    nel, nn = nelx * nely, (nelx + 1) * (nely + 1)
    g = torch.Generator(device="cpu").manual_seed(0)
    rnd = lambda *shape: torch.rand(*shape, generator=g, dtype=torch.float64).to(dev)  # noqa: E731
    x, hs = rnd(rows, nel), rnd(nel) + 1.0
    y = torch.empty_like(x)
    mesh = virtual_heat.ScalarMesh.build(nelx, nely, "bottom_edge", dev)
    sfield, w = rnd(rows, nn), rnd(rows, nel)
    X = torch.empty(rows, nel * 8, dtype=torch.float64, device=dev)
    F = torch.empty(rows, nn, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    h = stream.cuda_stream
    calls = {
        "density_filter_kernel": (
            lambda: lib.sttopt_density_filter(nelx, nely, rows, rmin, x.data_ptr(), nel,
                                              hs.data_ptr(), 1, y.data_ptr(), nel, h),
            rows * nel * 8 * 3),
        "heaviside_kernel": (
            lambda: lib.sttopt_heaviside(rows * nel, 16.0, 0.5, 0, x.data_ptr(), None,
                                         y.data_ptr(), h),
            rows * nel * 8 * 2),
        "gauss_gradients_kernel<true> (direction field)": (
            lambda: lib.sttopt_gauss_gradients(nelx, nely, rows, mesh._B_host,
                                               sfield.data_ptr(), nn, None, 0, 1.0, 1,
                                               X.data_ptr(), nel * 8, h),
            rows * (nn * 8 + nel * 64)),
        "divergence_load_kernel": (
            lambda: lib.sttopt_divergence_load(nelx, nely, rows, mesh._B_host, w.data_ptr(),
                                               nel, X.data_ptr(), nel * 8, 0.25,
                                               F.data_ptr(), nn, h),
            rows * (nel * 72 + nn * 8)),
    }
    out = {}
    # one MMA step on the space-time design (x and t: n = 2 nel variables, m = rows
    # constraints); wall clock of mmasub, whose Newton and line-search loops read two
    # numbers back per candidate
    from sttopt_amd import mma

    n, m = 2 * nel, rows
    xv = rnd(n) * 0.8 + 0.1
    g0, gv, G = rnd(n) - 0.5, rnd(m) - 0.5, (rnd(m, n) - 0.5) * 0.01
    move = torch.full_like(xv, 0.05)

    def mma_step(stats):
        return mma.unit_box_step(mma.History.initial(xv), 0, xv, move, 0.0, g0, gv, G,
                                 a0=1.0, c=1000.0, raa0=1e-5, stats=stats)

    mma_step({})
    torch.cuda.synchronize()
    stats = {}
    t0 = time.perf_counter()
    for _ in range(3):
        mma_step(stats)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / 3
    out["mma_unit_box_step"] = {"m": m, "n": n, "wall_ms": ms,
                                "newton_steps": stats.get("subsolv_newton_steps", 0) / 3,
                                "cap_hits": stats.get("subsolv_cap_hits", 0)}
    for name, (fn, nbytes) in calls.items():
        for _ in range(3):
            _lib.check(fn(), name)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(launches):
            fn()
        end.record(stream)
        end.synchronize()
        us = start.elapsed_time(end) * 1e3 / launches
        out[name] = {"avg_launch_us": us, "algorithmic_bytes_per_launch": nbytes,
                     "achieved_GBs": nbytes / us / 1e3,
                     "frac": nbytes / us / 1e3 / HBM_PEAK_GBS}
    print(json.dumps({"grid": f"{nelx}x{nely}", "rows": rows, "rmin": rmin,
                      "kernels": out}), flush=True)


def pmc_traffic():
    """HBM bytes per launch of the outer PCG's fused fp64 apply from the committed counter
    summary: FETCH_SIZE and WRITE_SIZE (KiB per launch, summed over that kernel's
    dispatches in one bench step and divided by its launches that still had rows), each
    scaled by the ratio of the bytes a plain copy moves to what the same counter reads
    for it.  None when the summary is absent."""
    try:
        with open(PMC_SUMMARY) as fh:
            kv = dict(line.split(":", 1) for line in fh if line.startswith("pmc."))
        kv = {k.strip(): float(v) for k, v in kv.items()}
        fetch = kv["pmc.apply_f64.fetch_kib"] * kv["pmc.copy.read_bytes"] / kv["pmc.copy.fetch_kib"]
        write = kv["pmc.apply_f64.write_kib"] * kv["pmc.copy.write_bytes"] / kv["pmc.copy.write_kib"]
        return fetch + write  # per launch with rows, as the live figure counts launches
    except (OSError, KeyError, ValueError, ZeroDivisionError):
        return None


def _time_apply(torch, _lib, args, stream, launches):
    lib = _lib.load()
    for _ in range(3):
        _lib.check(lib.sttopt_apply(*args), "sttopt_apply")
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(launches):
        lib.sttopt_apply(*args)
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / 1e3 / launches


def apply_roofline(torch, _lib, torch_fem, p, rows, dev, live, launches=50):
    """The fp64 K(rho)p apply of the outer PCG against its algorithmic bytes.  `live`:
    (seconds, bytes, launches) from HIP events around each of its launches in the timed
    steps -- the reported figure.  These are the launches from the second PCG iteration
    on, which carry the direction update (p_new = z + beta p_old) and, every other launch,
    the x updates of two iterations (x += alpha_prev p_prev, then x += alpha p_old): per
    row still iterating, p_old, z (fp32) and rho read and p_new and Ap written,
    8 (3 ndof + nel) + 4 ndof bytes, 8 (6 ndof + nel) + 4 ndof in the launches that read
    p_prev and read and write x; per row that retired the iteration before, 24 or 32 ndof
    (its x update only).  Also timed in a separate loop on all rows: the plain apply
    (8 (2 ndof + nel) bytes per row) and the public operator() with its input
    projection."""
    args, keep, stream = _apply_args(torch, _lib, p, rows, dev, torch.float64, _lib.F64)
    sec = _time_apply(torch, _lib, args, stream, launches)
    del keep
    args, keep, stream = _apply_args(torch, _lib, p, rows, dev, torch.float64, _lib.F64,
                                     _lib.MODE_OPERATOR)
    sec_op = _time_apply(torch, _lib, args, stream, launches)
    del keep
    nbytes = rows * 8 * (2 * p.ndof + p.nel)
    live_sec, live_bytes, live_n = live
    achieved = live_bytes / live_sec / 1e9
    return {
        "kernel": "stencil_tile_kernel<double,double,double,2,OPERATOR_DIR,DOT> (fp64 "
                  "K(rho)p of the outer PCG with the direction update p = z + beta p, x += "
                  "alpha p fused in, PCG iterations 2 on), HIP events around each launch in "
                  "the timed steps",
        "launches_in_timed_steps": int(live_n),
        "all_rows_loop": {"avg_launch_us": sec * 1e6, "achieved": nbytes / sec / 1e9,
                          "frac": nbytes / sec / 1e9 / HBM_PEAK_GBS,
                          "algorithmic_bytes_per_launch": nbytes},
        "operator_with_input_projection": {"avg_launch_us": sec_op * 1e6,
                                           "achieved": nbytes / sec_op / 1e9,
                                           "frac": nbytes / sec_op / 1e9 / HBM_PEAK_GBS},
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": pmc_traffic(),
        "algorithmic_bytes_per_launch": live_bytes / live_n,
        "avg_launch_us": live_sec / live_n * 1e6,
    }


def uneven_workload(torch, dev, n_instances=64):
    """The uneven batch as two callables that solve it with compliance gradients, both
    forming the densities and the compliance per grid: `batched(info=None)` (the
    instances, rows of their grid's density, in one batched MG-PCG,
    `uneven.solve_instances`) and `per_grid()` (one `femsolve` per grid)."""
    from sttopt_amd import compliance, fem, torch_fem, torch_solve, uneven, workloads

    specs = workloads.uneven_batch(n_instances, seed=0)
    KE = torch.as_tensor(fem.plane_stress_KE(0.3), dtype=torch.float64, device=dev)
    base = []
    for s in specs:
        mask = torch.zeros(s["loads"].shape[1], dtype=torch.bool, device=dev)
        mask[torch.as_tensor(s["freedofs"], device=dev)] = True
        edof = torch.as_tensor(fem.element_dof_map(s["nelx"], s["nely"]), device=dev)
        base.append((s, mask, edof, torch.as_tensor(s["loads"], dtype=torch.float64, device=dev)))

    def batched(info=None):
        # as per_grid below: the densities and the compliance formed per grid (one
        # stacked tensor each); the instances are rows of those, all solved together
        grids = {}
        for s, mask, edof, F in base:
            grids.setdefault((s["nelx"], s["nely"]), []).append((s, mask, edof, F))
        insts, dens_of = [], []
        for (nelx, nely), items in grids.items():
            x = torch.stack([torch.as_tensor(s["x"], dtype=torch.float64, device=dev).flatten()
                             for s, _, _, _ in items]).requires_grad_(True)
            dens = torch_fem.simp_density(x, 1e-9, 1.0, 3.0)
            dens_of.append((nelx, nely, dens, len(items)))
            for i, (s, mask, edof, F) in enumerate(items):
                insts.append(uneven.Instance(nelx, nely, dens[i], F, mask, edof, KE))
        Us = uneven.solve_instances(insts, info=info)
        total, start = 0, 0
        for nelx, nely, dens, n in dens_of:
            U = torch.cat(Us[start:start + n])
            start += n
            ce = compliance._StrainEnergy.apply(U, KE, nelx, nely)
            total = total + torch.sum(dens.repeat_interleave(2, dim=0) * ce)
        total.backward()

    def per_grid():
        grids = {}
        for s, mask, edof, F in base:
            grids.setdefault((s["nelx"], s["nely"]), []).append((s, mask, edof, F))
        for (nelx, nely), items in grids.items():
            x = torch.stack([torch.as_tensor(s["x"], dtype=torch.float64, device=dev).flatten()
                             for s, _, _, _ in items]).requires_grad_(True)
            dens = torch_fem.simp_density(x, 1e-9, 1.0, 3.0).repeat_interleave(2, dim=0)
            F = torch.cat([f for _, _, _, f in items])
            U = torch_solve.femsolve(dens, F, items[0][2], KE, items[0][1], nelx, nely)
            ce = compliance._StrainEnergy.apply(U, KE, nelx, nely)
            torch.sum(dens * ce).backward()

    def solver_only():
        """The same forward solves with the hierarchies built once, untimed: returns
        `(batched, per_grid)` callables that only solve (`uneven.batch_solve` over all
        grids against `torch_mg.native_mgpcg` once per grid)."""
        from sttopt_amd import torch_mg

        grids = {}
        for s, mask, edof, F in base:
            grids.setdefault((s["nelx"], s["nely"]), []).append((s, mask, edof, F))
        vcs, rhs = [], []
        for (nelx, nely), items in grids.items():
            x = torch.stack([torch.as_tensor(s["x"], dtype=torch.float64, device=dev).flatten()
                             for s, _, _, _ in items])
            dens = torch_fem.simp_density(x, 1e-9, 1.0, 3.0).repeat_interleave(2, dim=0)
            levels = torch_mg.build_hierarchy(dens, items[0][2], KE, items[0][1], nelx, nely)
            vcs.append(torch_mg.VCycle(levels, dtype=torch.float32))
            F = torch.cat([f for _, _, _, f in items])
            rhs.append(torch_fem.project(F, items[0][1]))

        def one_launch_per_pass():
            uneven.batch_solve(vcs, rhs, rtol=1e-8, max_iter=500)

        def one_call_per_grid():
            for vc, b in zip(vcs, rhs):
                torch_mg.native_mgpcg(vc, b, rtol=1e-8, max_iter=500)

        return one_launch_per_pass, one_call_per_grid

    batched.solver_only = solver_only
    return batched, per_grid


def uneven_throughput(torch, dev, n_instances=64, reps=3):
    """Solves with compliance gradients per second on the uneven batch: `n_instances`
    cantilever instances with grids drawn from {64x32, 128x64, 256x128, 512x256}, two
    load cases each (a top-edge traction and a seeded point load), seeded smooth densities
    at volume fraction 0.5; rtol 1e-8, fp32 V-cycles.  Timed: the whole batch in one
    batched MG-PCG (`uneven.solve_instances`) with its batched adjoint, against the same
    solves as one `femsolve` per grid (one grid per call, as the reference batches)."""
    batched, per_grid = uneven_workload(torch, dev, n_instances)
    solve_batched, solve_per_grid = batched.solver_only()
    out = {}
    for name, fn in (("batched", batched), ("per_grid", per_grid),
                     ("solve_batched", solve_batched), ("solve_per_grid", solve_per_grid)):
        fn()  # warm (hierarchy shapes, autotuned schedules)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0) / reps
    info = {}
    batched(info)
    solves = 2 * n_instances
    return {
        "workload": f"{n_instances} cantilever instances, grids 64x32..512x256, 2 load cases "
                    "each, rtol 1e-8, fp32 V-cycle, compliance gradients",
        "groups": info.get("groups"),
        "solves": solves,
        "solves_per_s": solves / out["batched"],
        "ms_per_batch": 1e3 * out["batched"],
        "per_grid_calls_solves_per_s": solves / out["per_grid"],
        "forward_solves_only": {
            "note": "hierarchies built once, untimed; forward MG-PCG only",
            "batched_ms": 1e3 * out["solve_batched"],
            "per_grid_calls_ms": 1e3 * out["solve_per_grid"],
        },
        "forward_iterations": info.get("forward_n_iter"),
        "adjoint_iterations": info.get("backward_n_iter"),
    }


def sequential_fp64(torch, p, rows, E):
    """The plain fp64 path one row per call: SIMP rows and loads as in the fused
    prologue, then FemSolve (fp64 V-cycle) + backward, one row at a time."""
    from sttopt_amd import compliance, torch_fem, torch_mg, torch_solve

    old = torch_mg.VCYCLE_DTYPE
    torch_mg.VCYCLE_DTYPE = torch.float64
    try:
        dev = p.xPhys.device
        mask = p.mask
        dens_rows, loads = [], []
        for k in range(rows):
            if k == 0:
                dens_rows.append(torch_fem.simp_density(p.xPhys, *E).flatten())
                loads.append(p.F)
            else:
                m = compliance.time_mask(p.tPhys, p.stage_times[k - 1], p.beta)
                xt = (p.xPhys * m).flatten()
                dens_rows.append(torch_fem.simp_density(xt, *E))
                F = torch.zeros(p.ndof, dtype=torch.float64, device=dev)
                C = torch_sparse_gravity(torch, p)
                F[1::2] = -(C @ xt)
                loads.append(F)

        def one(k):
            d = dens_rows[k].detach().clone().requires_grad_(True)
            U = torch_solve.femsolve(d, loads[k], p.edofMat, p.KE, mask, p.nelx, p.nely)
            ce = compliance._element_strain_energy(U, p.edofMat, p.KE, p.nely, p.nelx).reshape(-1)
            torch.sum(d * ce).backward()

        one(0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(rows):
            one(k)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        return {"rows_per_s": rows / sec, "seconds": sec}
    finally:
        torch_mg.VCYCLE_DTYPE = old


_C_CACHE = {}


def torch_sparse_gravity(torch, p):
    key = (p.nelx, p.nely)
    if key not in _C_CACHE:
        import numpy as np

        from sttopt_amd import fem

        C = fem.gravity_load_matrix(p.nelx, p.nely).tocsr()
        _C_CACHE[key] = torch.sparse_csr_tensor(
            torch.as_tensor(C.indptr), torch.as_tensor(C.indices), torch.as_tensor(C.data),
            size=C.shape, dtype=torch.float64, device=p.xPhys.device)
        assert np is not None
    return _C_CACHE[key]


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, E):
    """The oracle's torch-CPU fp64 restatement of the reference algorithm (index_add_
    matvec, einsum Galerkin, dense Cholesky, retiring PCG, alpha*U adjoint) on a
    bounded sample: the whole-structure row and one stage row of the same mesh,
    forward solve + adjoint + element gradient, the median of 5 runs after 2 warm-ups."""
    import numpy as np
    import torch

    from oracle import fem_direct, mgcg_port

    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    nelx, nely = p.nelx, p.nely
    edof = torch.as_tensor(fem_direct.element_dofs(nelx, nely))
    KE = torch.as_tensor(fem_direct.plane_stress_KE(0.3))
    mask = p.mask.cpu()
    x = p.xPhys.cpu().numpy()
    t = p.tPhys.cpu().numpy()
    m = fem_direct.time_mask(t, p.stage_times[len(p.stage_times) // 2], p.beta).ravel()
    xt = x.ravel() * m
    Fg = np.zeros(p.ndof)
    Fg[1::2] = -(fem_direct.gravity_matrix(nelx, nely) @ xt)
    dens = torch.as_tensor(np.stack([fem_direct.simp(x.ravel(), *E), fem_direct.simp(xt, *E)]))
    F = torch.stack([p.F.cpu(), torch.as_tensor(Fg)])
    sample_rows, repeats, warmups = 2, 5, 2
    times = []
    for i in range(warmups + repeats):
        t0 = time.perf_counter()
        U, n, levels = mgcg_port.mg_solve(F, dens, edof, KE, mask, nelx, nely, rtol=1e-8)
        Ue = U[..., edof]
        ce = torch.einsum("rei,ij,rej->re", Ue, KE, Ue)
        g = 2 * mgcg_port.matvec(U, dens, edof, KE, p.ndof)
        mgcg_port.adjoint(levels, U, F, g, edof, KE, rtol=1e-8)
        if i >= warmups:
            times.append(time.perf_counter() - t0)
        assert ce.shape[0] == sample_rows
    sec = float(np.median(times))
    return {
        "value": sample_rows / sec,
        "unit": "solves+grads/s",
        "cores": threads,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"{sample_rows} of the {1 + len(p.stage_times)} rows (whole structure + one "
                  f"stage) at {nelx}x{nely}: hierarchy + MG-PCG to 1e-8 ({n} it) + adjoint + "
                  f"gradient; median of {repeats} runs after {warmups} warm-ups "
                  f"({sum(times):.1f} s timed)",
        "seconds": sec,
    }


if __name__ == "__main__":
    main()

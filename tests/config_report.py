"""Per-configuration report of the solve path on the device: relative errors of U, c,
dc/drho and dc/dt against the fp64 direct oracle, PCG iterations per row, and the time
of one solve + gradient (median of 3 after a warm-up), for the five workload
configurations and the benchmark workload.

    python -u tests/config_report.py [--json out.json]

C1-C4 and the reduced C5 recipe are compared with the oracle at full size.  At 2048x1024
(C5) and 1024x512 (M) the direct oracle is too slow for every row: C5 reports the true
fp64 residual of every row and its speed-up over the plain per-stage fp64 path (fp64
V-cycle, one row per call); M reports the true residuals and the oracle's U and c of the
whole-structure row.  Prints one Markdown table and, with --json, writes the rows.
"""

from __future__ import annotations

import argparse
import json
import statistics
import sys
import time

import numpy as np

import conftest  # noqa: F401  (the package import)
from conftest import rel_err
from oracle import fem_direct
from test_gpu_configs import EMAX, EMIN, NU, PENAL, _heat_chain

import bench  # the repository root is on sys.path (conftest)


def log(msg: str) -> None:
    print(f"# {msg}", flush=True)


def timed(torch, fn, reps: int = 3) -> float:
    """Median seconds of `fn()` over `reps` runs after one warm-up."""
    fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        out.append(time.perf_counter() - t0)
    return statistics.median(out)


def _rows_info(info) -> str:
    it = info.get("row_iter_list") or []
    return f"{max(it)} / {statistics.mean(it):.1f}" if it else "-"


def _space_time_errors(p, c, stages, U, x, t):
    c_ref, dcx_ref, st_ref, U_ref = fem_direct.space_time(
        p.xPhys.cpu().numpy(), p.tPhys.cpu().numpy(), p.stage_times, p.beta,
        fem_direct.plane_stress_KE(NU), fem_direct.element_dofs(p.nelx, p.nely),
        p.F.cpu().numpy(), p.freedofs.cpu().numpy(), p.C)
    eu = max(rel_err(U[k].detach().cpu().numpy(), U_ref[k]) for k in range(U_ref.shape[0]))
    # the whole-structure row carries no load in C2 (c = 0 there)
    ec = max(([abs(float(c.detach()) - c_ref) / abs(c_ref)] if c_ref != 0 else [])
             + [abs(float(stages[k].detach()) - s[0]) / abs(s[0]) for k, s in enumerate(st_ref)])
    ex = rel_err(x.grad.cpu().numpy().ravel(), dcx_ref + sum(s[1] for s in st_ref))
    et = rel_err(t.grad.cpu().numpy().ravel(), sum(s[2] for s in st_ref))
    return eu, ec, ex, et


def _space_time_step(compliance, p):
    def run():
        x = p.xPhys.clone().requires_grad_(True)
        t = p.tPhys.clone().requires_grad_(True)
        c, stages, U = compliance.batched_whole_and_gravity_compliance(
            x, t, p.KE, p.edofMat, EMIN, EMAX, PENAL, p.freedofs, p.F, p.ndof, p.C, p.beta,
            p.stage_times)
        (c + sum(stages)).backward()
        return c, stages, U, x, t
    return run


def c1(torch, dev):
    from sttopt_amd import compliance, torch_fem, torch_mg, torch_solve, workloads

    p = workloads.mbb(device=dev)
    old = torch_mg.VCYCLE_DTYPE
    torch_mg.VCYCLE_DTYPE = torch.float64
    info = {}

    def run():
        x = p.xPhys.clone().requires_grad_(True)
        dens = torch_fem.simp_density(x.flatten(), EMIN, EMAX, PENAL)
        U = torch_solve.femsolve(dens, p.F, p.edofMat, p.KE, p.mask, p.nelx, p.nely,
                                 max_coarse_elements=200, info=info)
        ce = compliance._element_strain_energy(U, p.edofMat, p.KE, p.nely, p.nelx)
        c = torch.sum((EMIN + x**PENAL * (EMAX - EMIN)) * ce)
        c.backward()
        return c, U, x
    try:
        c, U, x = run()
        sec = timed(torch, run)
    finally:
        torch_mg.VCYCLE_DTYPE = old
    c_ref, dcx_ref, U_ref = fem_direct.whole_compliance(
        p.xPhys.cpu().numpy(), fem_direct.plane_stress_KE(NU),
        fem_direct.element_dofs(p.nelx, p.nely), p.F.cpu().numpy(), p.freedofs.cpu().numpy())
    return dict(config="C1", workload="MBB 120x40, 1 row", vcycle="fp64",
                iterations=f"{info['forward_n_iter']} (+{info['backward_n_iter']} adjoint)",
                err_u=rel_err(U.detach().cpu().numpy(), U_ref),
                err_c=abs(float(c.detach()) - c_ref) / c_ref,
                err_drho=rel_err(x.grad.cpu().numpy().ravel(), dcx_ref), err_dt=None,
                ms=1e3 * sec, solves_per_s=1 / sec)


def c2(torch, dev):
    from sttopt_amd import compliance, torch_mg, workloads

    p = workloads.space_time(240, 120, 10, beta=16.0, seed=0, point_load=False, device=dev)
    old = torch_mg.VCYCLE_DTYPE
    torch_mg.VCYCLE_DTYPE = torch.float64
    compliance.MAX_COARSE_ELEMENTS = 2048
    try:
        run = _space_time_step(compliance, p)
        c, stages, U, x, t = run()
        info = dict(compliance.LAST_SOLVE_INFO)
        sec = timed(torch, run)
    finally:
        torch_mg.VCYCLE_DTYPE = old
        compliance.MAX_COARSE_ELEMENTS = None
    eu, ec, ex, et = _space_time_errors(p, c, stages, U, x, t)
    rows = 1 + len(p.stage_times)
    return dict(config="C2", workload="space-time 240x120, 1 + 10 rows, beta 16 (coarsest 60x30)",
                vcycle="fp64", iterations=_rows_info(info), err_u=eu, err_c=ec, err_drho=ex,
                err_dt=et, ms=1e3 * sec, solves_per_s=rows / sec)


def c3(torch, dev):
    from sttopt_amd import workloads

    hp = workloads.heat(480, 240, seed=0, device=dev)
    ep = workloads.space_time(480, 240, 4, beta=16.0, seed=0, point_load=False, device=dev)

    def run():
        rho = hp.rho.reshape(240, 480).clone().requires_grad_(True)
        T, t, L = _heat_chain(hp, ep, rho, 1e-10)
        L.backward()
        return T, t, L, rho
    T, t, L, rho = run()
    sec = timed(torch, run)
    nelx, nely = 480, 240
    r = hp.rho.cpu().numpy()
    edof1 = fem_direct.element_dofs(nelx, nely, 1)
    Kh = fem_direct.assemble(fem_direct.diffusion_KE(), 1e-3 + r**3, edof1, (nelx + 1) * (nely + 1))
    fixed = hp.fixed.cpu().numpy()
    T_ref = fem_direct.lifted_solve(Kh, hp.F.cpu().numpy(), np.zeros(fixed.size), fixed)
    tbar = T_ref[edof1].mean(1)
    tmax = tbar.max()
    KE = fem_direct.plane_stress_KE(NU)
    edof2 = fem_direct.element_dofs(nelx, nely)
    dLdx, dLdt, Lref = np.zeros(nelx * nely), np.zeros(nelx * nely), 0.0
    for ti in ep.stage_times:
        cg, gx, gt, _ = fem_direct.gravity_compliance(
            r.reshape(nely, nelx), (tbar / tmax).reshape(nely, nelx), ti, ep.beta, KE, edof2,
            ep.freedofs.cpu().numpy(), ep.C)
        Lref, dLdx, dLdt = Lref + cg, dLdx + gx, dLdt + gt
    gT = np.zeros(fixed.size)
    np.add.at(gT, edof1.ravel(), np.repeat(dLdt / tmax / 4, 4))
    lam = fem_direct.lifted_solve(Kh, gT, np.zeros(fixed.size), fixed)
    dchi = -np.einsum("ei,ij,ej->e", lam[edof1], fem_direct.diffusion_KE(), T_ref[edof1])
    grad_ref = dLdx + dchi * 3 * r**2
    return dict(config="C3", workload="heat 480x240 -> time field -> 4-stage compliance",
                vcycle="fp32", iterations="rtol 1e-10 heat, 1e-8 elastic",
                err_u=rel_err(T.detach().cpu().numpy(), T_ref),
                err_c=abs(float(L.detach()) - Lref) / abs(Lref),
                err_drho=rel_err(rho.grad.cpu().numpy().ravel(), grad_ref), err_dt=None,
                ms=1e3 * sec, solves_per_s=6 / sec,
                note=f"U column: T; t rel err "
                     f"{rel_err(t.detach().cpu().numpy().ravel(), tbar / tmax):.1e}; solves: "
                     "1 heat + 5 elastic rows")


def c4(torch, dev):
    from sttopt_amd import compliance, torch_fem, uneven, workloads

    specs = workloads.uneven_batch(64, seed=0)
    KE = torch.tensor(fem_direct.plane_stress_KE(NU), dtype=torch.float64, device=dev)
    insts, xs = [], []
    for s in specs:
        mask = torch.zeros(s["loads"].shape[1], dtype=torch.bool, device=dev)
        mask[torch.as_tensor(s["freedofs"], device=dev)] = True
        x = torch.tensor(s["x"], dtype=torch.float64, device=dev).requires_grad_(True)
        xs.append(x)
        insts.append(uneven.Instance(
            s["nelx"], s["nely"], torch_fem.simp_density(x.flatten(), EMIN, EMAX, PENAL),
            torch.tensor(s["loads"], dtype=torch.float64, device=dev), mask,
            torch.tensor(fem_direct.element_dofs(s["nelx"], s["nely"]), device=dev), KE))
    info = {}
    Us = uneven.solve_instances(insts, info=info)
    total = 0
    for inst, U in zip(insts, Us):
        total = total + torch.sum(inst.density * compliance._StrainEnergy.apply(
            U, KE, inst.nelx, inst.nely))
    total.backward()
    eu = ec = ex = 0.0
    seen = set()
    for s, U, x in zip(specs, Us, xs):
        key = (s["nelx"], s["nely"])
        if key in seen:
            continue
        seen.add(key)
        edof = fem_direct.element_dofs(*key)
        K = fem_direct.assemble(fem_direct.plane_stress_KE(NU),
                                fem_direct.simp(s["x"].ravel(), EMIN, EMAX, PENAL), edof,
                                s["loads"].shape[1])
        ce = 0
        for k in range(2):
            want = fem_direct.solve(K, s["loads"][k], s["freedofs"])
            eu = max(eu, rel_err(U[k].detach().cpu().numpy(), want))
            c_k = float(s["loads"][k] @ want)
            got = float(torch.as_tensor(s["loads"][k], device=dev) @ U[k].detach())
            ec = max(ec, abs(got - c_k) / abs(c_k))
            ce = ce + fem_direct.strain_energy(want, edof, fem_direct.plane_stress_KE(NU))
        gref = -PENAL * (EMAX - EMIN) * s["x"].ravel() ** (PENAL - 1) * ce
        ex = max(ex, rel_err(x.grad.cpu().numpy().ravel(), gref))
    line = bench.uneven_throughput(torch, dev)
    return dict(config="C4", workload="64 cantilevers 64x32..512x256 x 2 load cases, one batched "
                "MG-PCG", vcycle="fp32",
                iterations=f"{info['forward_n_iter']} (+{info['backward_n_iter']} adjoint)",
                err_u=eu, err_c=ec, err_drho=ex, err_dt=None, ms=line["ms_per_batch"],
                solves_per_s=line["solves_per_s"],
                note=f"one femsolve per grid: {line['per_grid_calls_solves_per_s']:.0f} solves/s")


def c5_reduced(torch, dev):
    from sttopt_amd import compliance, workloads

    p = workloads.space_time(256, 128, 16, beta=32.0, seed=0, mean=0.45, std=0.05,
                             point_load=True, time_field="print", sigma=4.0, device=dev)
    run = _space_time_step(compliance, p)
    c, stages, U, x, t = run()
    info = dict(compliance.LAST_SOLVE_INFO)
    sec = timed(torch, run)
    eu, ec, ex, et = _space_time_errors(p, c, stages, U, x, t)
    return dict(config="C5 recipe at 256x128", workload="1 + 16 rows, beta 32, print order",
                vcycle="fp32", iterations=_rows_info(info), err_u=eu, err_c=ec, err_drho=ex,
                err_dt=et, ms=1e3 * sec, solves_per_s=17 / sec)


def _large(torch, dev, p, name, workload, oracle_row0):
    from sttopt_amd import compliance

    E = (EMIN, EMAX, PENAL)
    run = _space_time_step(compliance, p)
    c, stages, U, x, t = run()
    info = dict(compliance.LAST_SOLVE_INFO)
    sec = timed(torch, run)
    rows = 1 + len(p.stage_times)
    resid = bench.true_residuals(torch, p, U, E)
    log(f"{name}: {1e3 * sec:.1f} ms per step; plain per-stage fp64 path next")
    seq = bench.sequential_fp64(torch, p, rows, E)
    row = dict(config=name, workload=workload, vcycle="fp32", iterations=_rows_info(info),
               err_u=None, err_c=None, err_drho=None, err_dt=None, ms=1e3 * sec,
               solves_per_s=rows / sec, true_residual_max=max(resid),
               fp64_per_stage_solves_per_s=seq["rows_per_s"],
               speedup=rows / sec / seq["rows_per_s"])
    if oracle_row0:
        log(f"{name}: direct oracle of the whole-structure row (about a minute)")
        c_ref, _, U_ref = fem_direct.whole_compliance(
            p.xPhys.cpu().numpy(), fem_direct.plane_stress_KE(NU),
            fem_direct.element_dofs(p.nelx, p.nely), p.F.cpu().numpy(),
            p.freedofs.cpu().numpy())
        row["err_u"] = rel_err(U[0].detach().cpu().numpy(), U_ref)
        row["err_c"] = abs(float(c.detach()) - c_ref) / abs(c_ref)
    return row


def c5(torch, dev):
    from sttopt_amd import workloads

    p = workloads.space_time(2048, 1024, 16, beta=32.0, seed=0, mean=0.45, std=0.05,
                             point_load=True, time_field="print", sigma=32.0, device=dev)
    return _large(torch, dev, p, "C5", "space-time 2048x1024, 1 + 16 rows, beta 32", False)


def m(torch, dev):
    from sttopt_amd import workloads

    p = workloads.space_time(bench.NELX, bench.NELY, bench.NSTAGE, beta=bench.BETA, seed=0,
                             time_field="print", sigma=16.0, device=dev)
    return _large(torch, dev, p, "M", f"space-time {bench.NELX}x{bench.NELY}, "
                  f"1 + {bench.NSTAGE} rows, beta {bench.BETA:g}", True)


def _fmt(v):
    if v is None:
        return "-"
    return f"{v:.1e}" if isinstance(v, float) else str(v)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--json")
    ap.add_argument("--only", nargs="*")
    args = ap.parse_args()
    import torch

    dev = torch.device("cuda:0")
    out = []
    for name, fn in (("C1", c1), ("C2", c2), ("C3", c3), ("C4", c4), ("C5r", c5_reduced),
                     ("C5", c5), ("M", m)):
        if args.only and name not in args.only:
            continue
        log(f"{name} ...")
        row = fn(torch, dev)
        log(json.dumps(row))
        out.append(row)
    print("| config | workload | V-cycle | iterations (max / mean per row) | U | c | dc/drho "
          "| dc/dt | ms per solve+grad | solves/s |")
    print("|---|---|---|---|---|---|---|---|---|---|")
    for r in out:
        print(f"| {r['config']} | {r['workload']} | {r['vcycle']} | {r['iterations']} | "
              f"{_fmt(r['err_u'])} | {_fmt(r['err_c'])} | {_fmt(r['err_drho'])} | "
              f"{_fmt(r['err_dt'])} | {r['ms']:.1f} | {r['solves_per_s']:.1f} |")
    for r in out:
        extra = {k: v for k, v in r.items() if k in ("note", "true_residual_max",
                                                     "fp64_per_stage_solves_per_s", "speedup")}
        if extra:
            print(f"- {r['config']}: " + ", ".join(f"{k} {_fmt(v) if isinstance(v, float) else v}"
                                                   for k, v in extra.items()))
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(out, fh, indent=1)
    sys.stdout.flush()


if __name__ == "__main__":
    main()

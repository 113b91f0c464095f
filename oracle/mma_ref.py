"""ORACLE -- test infrastructure only (imported by tests/ as the checker; never by the
product).

NumPy restatement of the MMA subproblem (sttopt/mma.py:150-557, after Svanberg's
mmasub.m / subsolv.m): the moving asymptotes, the move-limited bounds, the separable
approximations p0, q0, P, Q, b, and the primal-dual interior-point Newton solve (both
eliminations, m < n and m >= n; the backtracking line search one halving at a time).

Pinned by tests/test_oracle.py: the KKT conditions of the barrier subproblem at the
returned point (both branches; the checks written from the subproblem's definition),
the asymptote update on the hand-picked case of the reference's tests
(tests/test_mma.py:62-117: every branch of the oscillation factor and each clamp), and
convergence of repeated mmasub calls to the closed-form optima of Svanberg's cantilever
beam and of a quadratic with a linear constraint (tests/test_mma_toy_problems.py).
"""

from __future__ import annotations

import numpy as np


def mmasub(m, n, iteration, xval, xmin, xmax, xold1, xold2, df0dx, fval, dfdx, low, upp,
           a0, a, c, d, asyinit=0.5, asyclamp_min=0.01, asyclamp_max=10.0, move=0.5,
           raa0=1e-5, albefa=0.1, asyincr=1.2, asydecr=0.7):
    xrange = xmax - xmin
    if iteration < 2:
        low = xval - asyinit * xrange
        upp = xval + asyinit * xrange
    else:
        zzz = (xval - xold1) * (xold1 - xold2)
        factor = np.where(zzz > 0, asyincr, np.where(zzz < 0, asydecr, 1.0))
        low = xval - factor * (xold1 - low)
        upp = xval + factor * (upp - xold1)
        low = np.minimum(np.maximum(low, xval - asyclamp_max * xrange), xval - asyclamp_min * xrange)
        upp = np.maximum(np.minimum(upp, xval + asyclamp_max * xrange), xval + asyclamp_min * xrange)
    alfa = np.maximum(np.maximum(low + albefa * (xval - low), xval - move * xrange), xmin)
    beta = np.minimum(np.minimum(upp - albefa * (upp - xval), xval + move * xrange), xmax)
    xmamiinv = 1.0 / np.maximum(xrange, 1e-5)
    ux2 = (upp - xval) ** 2
    xl2 = (xval - low) ** 2
    p0 = np.maximum(df0dx, 0)
    q0 = np.maximum(-df0dx, 0)
    pq0 = 0.001 * (p0 + q0) + raa0 * xmamiinv
    p0 = (p0 + pq0) * ux2
    q0 = (q0 + pq0) * xl2
    P = np.maximum(dfdx, 0)
    Q = np.maximum(-dfdx, 0)
    PQ = 0.001 * (P + Q) + raa0 * xmamiinv[None, :]
    P = (P + PQ) * ux2[None, :]
    Q = (Q + PQ) * xl2[None, :]
    b = P @ (1 / (upp - xval)) + Q @ (1 / (xval - low)) - fval
    sol = subsolv(m, n, 1e-7, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d)
    return (*sol, low, upp)


def _residual(v, epsi, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d):
    x, y, z, lam, xsi, eta, mu, zet, s = v
    ux1, xl1 = upp - x, x - low
    plam, qlam = p0 + P.T @ lam, q0 + Q.T @ lam
    gvec = P @ (1 / ux1) + Q @ (1 / xl1)
    return np.concatenate([
        plam / ux1**2 - qlam / xl1**2 - xsi + eta,
        c + d * y - mu - lam,
        [a0 - zet - a @ lam],
        gvec - a * z - y + s - b,
        xsi * (x - alfa) - epsi,
        eta * (beta - x) - epsi,
        mu * y - epsi,
        [zet * z - epsi],
        lam * s - epsi,
    ])


def subsolv(m, n, epsimin, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d):
    x = 0.5 * (alfa + beta)
    y = np.ones(m)
    z = 1.0
    lam = np.ones(m)
    xsi = np.maximum(1.0 / (x - alfa), 1.0)
    eta = np.maximum(1.0 / (beta - x), 1.0)
    mu = np.maximum(np.ones(m), 0.5 * c)
    zet = 1.0
    s = np.ones(m)
    epsi = 1.0
    prob = (low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d)
    while epsi > epsimin:
        r = _residual((x, y, z, lam, xsi, eta, mu, zet, s), epsi, *prob)
        rnorm, rmax = np.linalg.norm(r), np.abs(r).max()
        it = 0
        while rmax > 0.9 * epsi and it < 200:
            it += 1
            ux1, xl1 = upp - x, x - low
            plam, qlam = p0 + P.T @ lam, q0 + Q.T @ lam
            gvec = P @ (1 / ux1) + Q @ (1 / xl1)
            GG = P / ux1**2 - Q / xl1**2
            dpsidx = plam / ux1**2 - qlam / xl1**2
            delx = dpsidx - epsi / (x - alfa) + epsi / (beta - x)
            dely = c + d * y - lam - epsi / y
            delz = a0 - a @ lam - epsi / z
            dellam = gvec - a * z - y - b + epsi / lam
            diagx = 2 * (plam / ux1**3 + qlam / xl1**3) + xsi / (x - alfa) + eta / (beta - x)
            diagy = d + mu / y
            diaglamyi = s / lam + 1.0 / diagy
            if m < n:
                blam = dellam + dely / diagy - GG @ (delx / diagx)
                A = np.zeros((m + 1, m + 1))
                A[:m, :m] = np.diag(diaglamyi) + (GG / diagx) @ GG.T
                A[:m, m] = a
                A[m, :m] = a
                A[m, m] = -zet / z
                sol = np.linalg.solve(A, np.concatenate([blam, [delz]]))
                dlam, dz = sol[:m], sol[m]
                dx = -delx / diagx - (GG.T @ dlam) / diagx
            else:
                dellamyi = dellam + dely / diagy
                A = np.zeros((n + 1, n + 1))
                A[:n, :n] = np.diag(diagx) + (GG.T / diaglamyi) @ GG
                A[:n, n] = -GG.T @ (a / diaglamyi)
                A[n, :n] = A[:n, n]
                A[n, n] = zet / z + a @ (a / diaglamyi)
                rhs = np.concatenate([-(delx + GG.T @ (dellamyi / diaglamyi)),
                                      [-(delz - a @ (dellamyi / diaglamyi))]])
                sol = np.linalg.solve(A, rhs)
                dx, dz = sol[:n], sol[n]
                dlam = (GG @ dx) / diaglamyi - dz * (a / diaglamyi) + dellamyi / diaglamyi
            dy = -dely / diagy + dlam / diagy
            dxsi = -xsi + epsi / (x - alfa) - (xsi * dx) / (x - alfa)
            deta = -eta + epsi / (beta - x) + (eta * dx) / (beta - x)
            dmu = -mu + epsi / y - (mu * dy) / y
            dzet = -zet + epsi / z - zet * dz / z
            ds = -s + epsi / lam - (s * dlam) / lam
            xx = np.concatenate([y, [z], lam, xsi, eta, mu, [zet], s])
            dxx = np.concatenate([dy, [dz], dlam, dxsi, deta, dmu, [dzet], ds])
            steg = 1.0 / max((-1.01 * dx / (x - alfa)).max(), (1.01 * dx / (beta - x)).max(),
                             (-1.01 * dxx / xx).max(), 1.0)
            old = (x, y, z, lam, xsi, eta, mu, zet, s)
            dirs = (dx, dy, dz, dlam, dxsi, deta, dmu, dzet, ds)
            for j in range(50):
                t = steg / 2.0**j
                cand = tuple(o + t * dv for o, dv in zip(old, dirs))
                r = _residual(cand, epsi, *prob)
                new = np.linalg.norm(r)
                if not new > rnorm:
                    break
            x, y, z, lam, xsi, eta, mu, zet, s = cand
            rnorm, rmax = new, np.abs(r).max()
        epsi *= 0.1
    return x, y, z, lam, xsi, eta, mu, zet, s


def kkt_residuals(sol, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d):
    """Stationarity / feasibility residuals (max-abs) of the subproblem at `sol`, and
    the complementarity products, from the subproblem's definition."""
    x, y, z, lam, xsi, eta, mu, zet, s = (np.asarray(v, float) for v in sol)
    ux1, xl1 = upp - x, x - low
    plam, qlam = p0 + P.T @ lam, q0 + Q.T @ lam
    stat = [plam / ux1**2 - qlam / xl1**2 - xsi + eta, c + d * y - mu - lam,
            np.atleast_1d(a0 - zet - a @ lam),
            P @ (1 / ux1) + Q @ (1 / xl1) - a * z - y + s - b]
    comp = [xsi * (x - alfa), eta * (beta - x), mu * y, lam * s, np.atleast_1d(zet * z)]
    return max(np.abs(r).max() for r in stat), comp

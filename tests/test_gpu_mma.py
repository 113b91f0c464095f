"""The MMA subproblem on the device against the NumPy oracle (oracle/mma_ref.py): the
subsolv point and multipliers on both eliminations (m < n, m >= n) and at a size where
the Gram product splits over many column chunks, the KKT conditions of the returned
point, mmasub's asymptotes and subproblem data, and the closed-form optima of the
reference's toy problems (sttopt/mma.py:150-557, tests/test_mma*.py)."""

import math
import warnings

import numpy as np
import pytest
import torch

from oracle import mma_ref

pytestmark = pytest.mark.gpu

KEYS = ("low", "upp", "alfa", "beta", "p0", "q0", "P", "Q", "a0", "a", "b", "c", "d")


def _problem(m, n, seed):
    rng = np.random.default_rng(seed)
    pr = dict(low=np.full(n, -1.0), upp=np.full(n, 1.0), alfa=np.full(n, -0.9),
              beta=np.full(n, 0.9), p0=rng.uniform(0.5, 2.0, n), q0=rng.uniform(0.5, 2.0, n),
              P=rng.uniform(0.1, 1.0, (m, n)), Q=rng.uniform(0.1, 1.0, (m, n)), a0=1.0,
              a=rng.uniform(0.1, 0.3, m), c=np.full(m, 10.0), d=rng.uniform(0.25, 1.0, m))
    pr["b"] = 0.8 * (pr["P"].sum(1) + pr["Q"].sum(1))
    return pr


def _t(a, dev):
    return torch.as_tensor(np.asarray(a), dtype=torch.float64, device=dev)


@pytest.mark.parametrize("m, n", [(3, 8), (3, 2), (5, 40), (17, 5000), (40, 70000), (100, 3000),
                                  (20, 12)])
def test_subsolv_matches_oracle_and_satisfies_kkt(dev, m, n):
    from sttopt_amd import mma

    pr = _problem(m, n, 100 * m + n)
    args = [_t(pr[k], dev) if k != "a0" else pr[k] for k in KEYS]
    sol = mma.subsolv(m, n, 1e-7, *args)
    got = [v.cpu().numpy() for v in sol]
    stat, comp = mma_ref.kkt_residuals(got, **{k: pr[k] for k in KEYS})
    assert stat < 1e-5, stat
    for prod in comp:
        assert np.all(np.abs(prod - 1e-6) < 1e-5)
    want = mma_ref.subsolv(m, n, 1e-7, *(pr[k] for k in KEYS))
    for g, w, name in zip(got, want, ("x", "y", "z", "lam", "xsi", "eta", "mu", "zet", "s")):
        np.testing.assert_allclose(g, w, rtol=1e-6, atol=1e-8, err_msg=name)
    assert got[3].max() > 1e-3


def test_mmasub_matches_oracle_on_first_and_later_iterations(dev):
    """Asymptotes, bounds and the subproblem's solution at iteration 0 (asymptotes from
    asyinit) and at iteration 3 with the reference's hand-picked oscillation and clamp
    cases (tests/test_mma.py:62-117)."""
    from sttopt_amd import mma

    n, m = 6, 1
    xval = np.full(n, 5.0)
    xold1 = np.array([4.0, 4.0, 5.0, 4.0, 4.0, 4.0])
    xold2 = np.array([3.0, 5.0, 2.0, 3.0, 3.0, 4.0])
    low = np.array([2.0, 2.0, 2.0, -1000.0, 2.0, 4.8])
    upp = np.array([8.0, 8.0, 8.0, 8.0, 1000.0, 4.0])
    rng = np.random.default_rng(4)
    df0dx = rng.standard_normal(n)
    dfdx = rng.standard_normal((m, n))
    fval = np.array([0.3])
    for it in (0, 3):
        res = mma.mmasub(m, n, it, _t(xval, dev), _t(np.zeros(n), dev), _t(np.full(n, 10.0), dev),
                         _t(xold1, dev), _t(xold2, dev), 0.0, _t(df0dx, dev), _t(fval, dev),
                         _t(dfdx, dev), _t(low, dev), _t(upp, dev), 1.0, _t(np.zeros(m), dev),
                         _t(np.full(m, 2500.0), dev), _t(np.zeros(m), dev))
        want = mma_ref.mmasub(m, n, it, xval, np.zeros(n), np.full(n, 10.0), xold1, xold2, df0dx,
                              fval, dfdx, low, upp, 1.0, np.zeros(m), np.full(m, 2500.0), np.zeros(m))
        for g, w in zip(res, want):
            np.testing.assert_allclose(np.asarray(g.cpu()), w, rtol=1e-7, atol=1e-9)
        if it == 3:
            np.testing.assert_allclose(res[-2].cpu().numpy(), [2.6, 3.6, 2.0, -95.0, 2.6, 4.9], atol=1e-12)
            np.testing.assert_allclose(res[-1].cpu().numpy(), [9.8, 7.8, 8.0, 9.8, 105.0, 5.1], atol=1e-12)


@pytest.mark.parametrize("beam", [True, False])
def test_mmasub_converges_to_closed_form_optima(dev, beam):
    """Svanberg's cantilever beam and the quadratic with a linear constraint, driven
    through the device mmasub (no subsolv warning may fire)."""
    from sttopt_amd import mma

    def run(m, n, it, x, xmin, xmax, xold1, xold2, g0, fv, g, low, upp, a0, a, c, d):
        out = mma.mmasub(m, n, it, _t(x, dev), _t(xmin, dev), _t(xmax, dev), _t(xold1, dev),
                         _t(xold2, dev), 0.0, _t(g0, dev), _t(fv, dev), _t(g, dev), _t(low, dev),
                         _t(upp, dev), a0, _t(a, dev), _t(c, dev), _t(d, dev))
        return [v.cpu().numpy() for v in out]

    from test_oracle import _mma_toy

    with warnings.catch_warnings():
        warnings.simplefilter("error")
        x = _mma_toy(run, beam)
    if beam:
        root4 = np.array([61.0, 37.0, 19.0, 7.0, 1.0]) ** 0.25
        np.testing.assert_allclose(x, np.sum(root4) ** (1 / 3) * root4, rtol=1e-4)
    else:
        np.testing.assert_allclose(x, [1.0, 1.0], atol=1e-6)


def test_unit_box_step_and_floor_share(dev):
    from sttopt_amd import mma

    rng = np.random.default_rng(8)
    n, m = 300, 4
    x = _t(rng.uniform(0.2, 0.8, n), dev)
    hist = mma.History.initial(x)
    move = torch.full_like(x, 0.05)
    xmma, lam, hist2 = mma.unit_box_step(hist, 0, x, move, 0.0, _t(rng.standard_normal(n), dev),
                                         _t(rng.standard_normal(m), dev), _t(rng.standard_normal((m, n)), dev),
                                         a0=1.0, c=1000.0, raa0=1e-5)
    assert torch.all((xmma - x).abs() <= 0.05 + 1e-12)
    assert xmma.shape == (n,) and lam.shape == (m,)
    share = mma.floor_share(hist2, move, torch.ones(n, dtype=torch.bool, device=dev))
    assert 0.0 <= share <= 1.0 and math.isfinite(share)

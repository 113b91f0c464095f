"""Density filter, symmetric sparse products and the Heaviside projection on the device
against the oracle (oracle/filters_ref.py) and SciPy: the matrix-free stencil (register
windows for radii 1-4, the LDS tile up to 8, the direct kernel beyond) and the CSR
product, forward and adjoint
(sttopt/filters.py:67-76,108-131, torch_util.py:157-188)."""

import math

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import filters_ref

pytestmark = pytest.mark.gpu


def _t(a, dev):
    return torch.as_tensor(np.asarray(a), dtype=torch.float64, device=dev)


@pytest.mark.parametrize("nelx, nely, rmin", [(7, 5, 1.0), (7, 5, 1.5), (23, 17, 2.3), (70, 33, 4.5),
                                              (45, 21, 3.5), (40, 26, 6.2), (31, 29, 9.7),
                                              (130, 70, 3.0)])
def test_apply_density_filter_matches_oracle(dev, nelx, nely, rmin):
    from sttopt_amd import filters, torch_util

    H, Hs = filters.density_filter(nelx, nely, rmin)
    if nelx * nely <= 2500:
        Ho, Hso = filters_ref.filter_matrix(nelx, nely, rmin)
    else:  # the host matrix equals the oracle's on every smaller case (test_host)
        Ho, Hso = H.toarray(), Hs
    Ht = torch_util.symmetric_csr_to_tensor(H, dev, torch.float64)
    Hst = _t(Hs, dev)
    rng = np.random.default_rng(nelx + nely)
    x = rng.uniform(0, 1, (nely, nelx))
    xt = _t(x, dev).requires_grad_(True)
    y = filters.apply_density_filter(xt, Ht, Hst)
    want = filters_ref.apply(x, Ho, Hso)
    assert y.shape == (nely, nelx)
    np.testing.assert_allclose(y.detach().cpu().numpy(), want, rtol=1e-14, atol=1e-15)
    g = rng.standard_normal((nely, nelx))
    y.backward(_t(g, dev))
    np.testing.assert_allclose(xt.grad.cpu().numpy().ravel(), Ho.T @ (g.ravel() / Hso), rtol=1e-13, atol=1e-14)
    # the same matrix without its stencil tag goes through the CSR product
    plain = torch_util.csr_to_tensor(H, dev, torch.float64)
    assert torch_util.stencil_of(plain) is None
    y2 = filters.apply_density_filter(_t(x, dev), plain, Hst)
    np.testing.assert_allclose(y2.cpu().numpy(), want, rtol=1e-14, atol=1e-15)


def test_symmetric_matmul_general_csr_and_columns(dev):
    """A random symmetric sparse matrix (not a filter): (n,) and (n, k) operands, the
    backward equal to the same product."""
    from sttopt_amd import torch_util

    rng = np.random.default_rng(1)
    n = 300
    A = sp.random(n, n, density=0.03, random_state=2, format="csr")
    A = (A + A.T).tocsr()
    At = torch_util.symmetric_csr_to_tensor(A, dev, torch.float64)
    v = rng.standard_normal(n)
    V = rng.standard_normal((n, 4))
    np.testing.assert_allclose(torch_util.symmetric_matmul(At, _t(v, dev)).cpu().numpy(), A @ v, rtol=1e-13, atol=1e-13)
    Vt = _t(V, dev).requires_grad_(True)
    out = torch_util.symmetric_matmul(At, Vt)
    np.testing.assert_allclose(out.detach().cpu().numpy(), A @ V, rtol=1e-13, atol=1e-13)
    G = rng.standard_normal((n, 4))
    out.backward(_t(G, dev))
    np.testing.assert_allclose(Vt.grad.cpu().numpy(), A.T @ G, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("beta, eta", [(1.0, 0.5), (8.0, 0.3), (64.0, 0.5), (math.inf, 0.5)])
def test_heaviside_projection_and_derivative_match_oracle(dev, beta, eta):
    from sttopt_amd import filters

    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-0.1, 1.1, 999), [eta, 0.0, 1.0]]).reshape(-1, 2)
    xt = _t(x, dev).requires_grad_(True)
    y = filters.heaviside_projection(xt, beta, eta)
    np.testing.assert_allclose(y.detach().cpu().numpy(), filters_ref.heaviside(x, beta, eta), rtol=1e-14, atol=1e-15)
    g = rng.standard_normal(x.shape)
    y.backward(_t(g, dev))
    if math.isinf(beta):
        assert torch.all(xt.grad == 0)
        return
    d = filters_ref.heaviside_derivative(x, beta, eta)
    # 1 - tanh^2 cancels where the derivative is small: absolute error ~ eps * beta
    atol = 4 * np.finfo(float).eps * beta
    np.testing.assert_allclose(xt.grad.cpu().numpy(), d * g, rtol=1e-13, atol=atol * np.abs(g).max())
    np.testing.assert_allclose(filters.heaviside_projection_derivative(_t(x, dev), beta, eta).cpu().numpy(), d,
                               rtol=1e-13, atol=atol)

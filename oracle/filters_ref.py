"""ORACLE -- test infrastructure only (imported by tests/ as the checker; never by the
product).

Dense NumPy restatement of the density filter and the Heaviside projection
(sttopt/filters.py:38-76 and :108-131):

- filter_matrix: H[e1, e2] = max(0, rmin - |e1 - e2|) for every pair of elements of the
  nelx x nely grid whose offsets satisfy |di|, |dj| <= ceil(rmin) - 1, by an explicit
  loop over element pairs; Hs = row sums; apply: H x / Hs;
- heaviside / heaviside_derivative: the smoothed threshold and its derivative, the step
  at beta = inf.

Pinned by tests/test_oracle.py: symmetry and row sums, rmin <= 1 giving the identity
filter, an interior row at rmin = 1.5 written out by hand (1.5, four 0.5, four
1.5 - sqrt 2) and a corner's normalizer, constants preserved; the projection's fixed
points H(0) = 0, H(1) = 1, its small-beta limit (the identity), its derivative against
central differences, and the step.
"""

from __future__ import annotations

import math

import numpy as np


def filter_matrix(nelx: int, nely: int, rmin: float) -> tuple[np.ndarray, np.ndarray]:
    n = nelx * nely
    R = int(math.ceil(rmin)) - 1
    H = np.zeros((n, n))
    for e1 in range(n):
        i1, j1 = e1 % nelx, e1 // nelx
        for e2 in range(n):
            i2, j2 = e2 % nelx, e2 // nelx
            di, dj = i2 - i1, j2 - j1
            if abs(di) <= R and abs(dj) <= R:
                H[e1, e2] = max(0.0, rmin - math.sqrt(di * di + dj * dj))
    return H, H.sum(axis=1)


def apply(x: np.ndarray, H: np.ndarray, Hs: np.ndarray) -> np.ndarray:
    return (H @ x.ravel() / Hs).reshape(x.shape)


def heaviside(x: np.ndarray, beta: float, eta: float) -> np.ndarray:
    if math.isinf(beta):
        return np.where(x > eta, 1.0, np.where(x < eta, 0.0, 0.5))
    den = math.tanh(beta * eta) + math.tanh(beta * (1 - eta))
    return (math.tanh(beta * eta) + np.tanh(beta * (x - eta))) / den


def heaviside_derivative(x: np.ndarray, beta: float, eta: float) -> np.ndarray:
    den = math.tanh(beta * eta) + math.tanh(beta * (1 - eta))
    return beta * (1 - np.tanh(beta * (x - eta)) ** 2) / den

"""Shared test setup: the package import, the `gpu` marker, small helpers."""

from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

__graft_entry__.package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def dev():
    """cuda:0; a gpu test fails (does not skip) without a device."""
    import torch

    assert torch.cuda.is_available(), "gpu tests need a ROCm device"
    return torch.device("cuda:0")


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def density_fields(nelx, nely, rng):
    """Uniform, random and near-void density fields (the reference's protocol)."""
    nel = nelx * nely
    uniform = np.full(nel, 0.4)
    random = rng.uniform(0.0, 1.0, nel)
    near_void = random.copy()
    near_void[: nel // 3] = 1e-6
    return {"uniform": uniform, "random": random, "near_void": near_void}


def binary_design(nelx, nely):
    """Connected 0/1 cantilever: chords, clamped edge and a diagonal brace."""
    rows, cols = np.mgrid[0:nely, 0:nelx]
    x = np.zeros((nely, nelx))
    x[np.abs(rows / max(nely - 1, 1) - cols / max(nelx - 1, 1)) < 0.15] = 1.0
    x[:2, :] = 1.0
    x[-2:, :] = 1.0
    x[:, :2] = 1.0
    return x.ravel()

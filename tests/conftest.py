"""Shared test set-up: import paths, the gpu marker, and cached workloads."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "legacy-mars-global-climate-model_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def rel_err(a: np.ndarray, b: np.ndarray, scale: np.ndarray | None = None) -> float:
    """max |a-b| / max(|b|, s) with s a per-row scale (max |b| over the row)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if scale is None:
        if b.ndim >= 2:
            scale = np.max(np.abs(b.reshape(b.shape[0], -1)), axis=1).reshape((-1,) + (1,) * (b.ndim - 1))
        else:
            scale = np.abs(b)
    den = np.maximum(np.abs(b), np.maximum(scale, 1e-300))
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()

"""Shared test setup.

Markers: `gpu` tests need a ROCm device and the built libmagforce.so; the
rest run on the CPU (oracle, generators, host logic, ABI symbol table).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "magnetic-force_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx():
    import torch
    import magforce
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    c = magforce.Context(0)
    yield c
    c.close()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def max_rel_err(a, b):
    """max_i |a_i - b_i| / RMS(|b|) over rows."""
    a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(b, dtype=np.float64).reshape(-1, 3)
    rms = np.sqrt((b ** 2).sum(1).mean())
    return float(np.abs(a - b).max() / max(rms, 1e-300))


def jittered_block(nx, ny, nz, seed=0, jitter=0.1):
    from magforce import configs
    return configs.jittered_lattice(nx, ny, nz, seed=seed, jitter=jitter).astype(np.float64)


def random_moments(n, seed=1, scale=0.5):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 3)) * scale

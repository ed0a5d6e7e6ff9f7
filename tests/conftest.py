import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libphl.so on the device)")


def gpu_available():
    try:
        import torch  # noqa: F401  (only to count devices; no allocation)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def perturbed_state(n_total, n_cell, prim_inf, seed, amp=0.05):
    """Freestream plus a seeded smooth-looking perturbation on real cells."""
    rng = np.random.default_rng(seed)
    q = np.tile(np.asarray(prim_inf, dtype=np.float64)[:, None], (1, n_total))
    scale = np.array([1.0, 1.0, 1.0, 1.0, prim_inf[4]])
    noise = rng.uniform(-amp, amp, size=(5, n_cell)) * scale[:, None]
    q[:, :n_cell] += noise
    return q

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU-side oracle computation")


def gaussian_bunch(N, seed, sigma_r=(1e-3, 1e-3, 1e-3), sigma_p=(1e-3, 1e-3, 1e-3), p0=0.046181,
                   q_total=1e-9, halo_frac=0.0, halo_radius=5.0):
    """Seeded 6D Gaussian bunch (OPAL units: m, beta*gamma, C)."""
    rng = np.random.default_rng(seed)
    R = rng.normal(size=(N, 3)) * np.asarray(sigma_r)
    if halo_frac > 0:
        m = rng.random(N) < halo_frac
        d = rng.normal(size=(int(m.sum()), 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        d *= (halo_radius * rng.random(len(d)) ** (1 / 3))[:, None]
        R[m] = d * np.asarray(sigma_r)
    P = rng.normal(size=(N, 3)) * np.asarray(sigma_p)
    P[:, 2] += p0
    Q = np.full(N, q_total / N)
    return R, P, Q


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)

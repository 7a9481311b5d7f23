"""Shared test helpers: seeded inputs and device transfers."""
import numpy as np

from filsim_amd import params


def random_segments(fp, n, seed):
    rng = np.random.default_rng(seed)
    L = np.array(params.box(fp))
    Y = rng.uniform(0, 1, (n, 3)) * L
    F = rng.standard_normal((n, 3))
    T = rng.standard_normal((n, 3))
    return Y, F, T


def to_dev(torch, a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64,
                        device="cuda")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) /
                 max(np.linalg.norm(np.asarray(b)), 1e-300))

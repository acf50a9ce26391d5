"""Seeded synthetic point clouds and right-hand sides of the benchmark
configurations (all generated on the host in FP64 with numpy's default_rng)."""
from __future__ import annotations

import numpy as np


def uniform(n: int, dim: int, seed: int = 0) -> np.ndarray:
    """n points uniform in [0,1]^dim, as a (dim, n) Fortran-ordered array."""
    return np.asfortranarray(np.random.default_rng(seed).random((dim, n)))


def gaussian_mixture(n: int, dim: int = 3, components: int = 16, sigma: float = 0.02,
                     seed: int = 0) -> np.ndarray:
    """Mixture of isotropic Gaussians (equal weights, centres uniform in [0,1]^dim):
    highly uneven leaves for the same cardinality-balanced tree."""
    rng = np.random.default_rng(seed)
    centres = rng.random((components, dim))
    which = rng.integers(0, components, n)
    P = centres[which].T + sigma * rng.standard_normal((dim, n))
    return np.asfortranarray(P)


def normal_rhs(n: int, nrhs: int, seed: int = 1) -> np.ndarray:
    return np.asfortranarray(np.random.default_rng(seed).standard_normal((n, nrhs)))


def smooth_rhs(P: np.ndarray, nrhs: int, noise: float = 0.01, seed: int = 2) -> np.ndarray:
    """y = sin(2 pi x1) cos(2 pi x2) + noise * N(0,1), one column per rhs."""
    rng = np.random.default_rng(seed)
    base = np.sin(2 * np.pi * P[0]) * np.cos(2 * np.pi * P[1])
    return np.asfortranarray(base[:, None] + noise * rng.standard_normal((P.shape[1], nrhs)))

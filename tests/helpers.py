"""Shared helpers for the test-suite (seeded inputs, error norms)."""

from __future__ import annotations

import numpy as np
import torch


def nrel(a, b) -> float:
    """Norm-wise relative error ||a - b|| / ||b|| (in float64)."""
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().to(torch.complex128 if a.is_complex() else torch.float64).numpy()
    if isinstance(b, torch.Tensor):
        b = b.detach().cpu().to(torch.complex128 if b.is_complex() else torch.float64).numpy()
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def crandn(rng, *shape) -> np.ndarray:
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def smooth_object(rng, n: int, sigma: float = 3.0) -> np.ndarray:
    """Smooth random complex object (amplitude in [0.5, 1], phase in [-pi/2, pi/2])."""
    from scipy.ndimage import gaussian_filter

    a = gaussian_filter(rng.uniform(0.5, 1.0, (n, n)), sigma, mode="wrap")
    ph = gaussian_filter(rng.uniform(-np.pi / 2, np.pi / 2, (n, n)), sigma, mode="wrap")
    a = 0.5 + 0.5 * (a - a.min()) / max(a.max() - a.min(), 1e-12)
    ph = -np.pi / 2 + np.pi * (ph - ph.min()) / max(ph.max() - ph.min(), 1e-12)
    return (a * np.exp(1j * ph)).astype(np.complex64)


def disc_probe(n: int, radius: float, defocus: float = 400.0) -> np.ndarray:
    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    r2 = xx**2 + yy**2
    return ((r2 <= radius**2) * np.exp(1j * np.pi * r2 / defocus)).astype(np.complex64)


def raster_positions(rng, n_side: int, step: int, origin: int, jitter: int) -> np.ndarray:
    g = np.arange(n_side) * step + origin
    yy, xx = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([xx.ravel(), yy.ravel()], axis=1)
    pos = pos + rng.integers(-jitter, jitter + 1, pos.shape)
    return pos.astype(np.int64)


def to_dev(x, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)

"""Synthetic datasets for the benchmark, the serving mix and the tests.

Recipes: smooth random complex objects (uniform amplitude / phase noise,
Gaussian-smoothed with wrap-around, then affinely rescaled back to the stated
ranges), disc probes with a quadratic defocus phase, jittered raster scans and
Poisson-noised far-field intensities whose measured amplitude is
``sqrt(counts / scale)``.  Host-side generation in FP64 with NumPy
``default_rng(seed)``; far fields are computed on the device with torch.fft.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch


def _rescale(x: np.ndarray, lo: float, hi: float) -> np.ndarray:
    mn, mx = x.min(), x.max()
    return lo + (hi - lo) * (x - mn) / max(mx - mn, 1e-300)


def smooth_complex_object(rng, n: int, sigma: float, amp=(0.5, 1.0), phase=(-np.pi / 2, np.pi / 2)) -> np.ndarray:
    from scipy.ndimage import gaussian_filter

    a = _rescale(gaussian_filter(rng.uniform(*amp, (n, n)), sigma, mode="wrap"), *amp)
    p = _rescale(gaussian_filter(rng.uniform(*phase, (n, n)), sigma, mode="wrap"), *phase)
    return (a * np.exp(1j * p)).astype(np.complex64)


def defocus_disc_probe(n: int, radius: float, defocus_div: float = 400.0) -> np.ndarray:
    """Circular aperture of ``radius`` px times ``exp(i pi r^2 / defocus_div)``."""
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    r2 = (yy - n / 2) ** 2 + (xx - n / 2) ** 2
    return ((r2 <= radius * radius) * np.exp(1j * np.pi * r2 / defocus_div)).astype(np.complex64)


def jittered_raster(rng, n_side: int, step: int, origin: int, jitter: int) -> np.ndarray:
    g = origin + step * np.arange(n_side)
    yy, xx = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([xx.ravel(), yy.ravel()], axis=1)
    pos = pos + rng.integers(-jitter, jitter + 1, pos.shape)
    return pos.astype(np.int64)


@dataclass
class ConventionalDataset:
    sample: torch.Tensor       # [H, W] complex64 (ground truth)
    probe: torch.Tensor        # [h, w] complex64
    positions: torch.Tensor    # [P, 2] int64 (x, y)
    measured: torch.Tensor     # [P, h, w] float32 amplitudes


@torch.no_grad()
def conventional_dataset(object_n: int, probe_n: int, probe_radius: float, n_side: int, step: int, origin: int,
                         jitter: int, photons: float, sigma: float, device, seed: int = 0,
                         chunk: int = 64) -> ConventionalDataset:
    """Conventional ptychography data set: far-field amplitudes of every scan
    position, Poisson noise at ``photons`` counts at the brightest pixel."""
    rng = np.random.default_rng(seed)
    obj = smooth_complex_object(rng, object_n, sigma)
    probe = defocus_disc_probe(probe_n, probe_radius)
    pos = jittered_raster(rng, n_side, step, origin, jitter)
    if pos.min() < 0 or pos.max() + probe_n > object_n:
        raise ValueError("scan leaves the object")
    dev = torch.device(device)
    s = torch.from_numpy(obj).to(dev)
    p = torch.from_numpy(probe).to(dev)
    pt = torch.from_numpy(pos).to(dev)
    rows = torch.arange(probe_n, device=dev)
    inten = torch.empty((len(pos), probe_n, probe_n), dtype=torch.float32, device=dev)
    for k in range(0, len(pos), chunk):
        pc = pt[k : k + chunk]
        idx = (pc[:, 1, None, None] + rows[None, :, None]) * object_n + pc[:, 0, None, None] + rows[None, None, :]
        ff = torch.fft.fftshift(torch.fft.fft2(s.reshape(-1)[idx] * p), dim=(-2, -1))
        inten[k : k + chunk] = ff.abs().square()
    scale = photons / float(inten.max())
    gen = torch.Generator(device=dev).manual_seed(seed)
    counts = torch.poisson(inten * scale, generator=gen)
    measured = torch.sqrt(counts / scale)
    return ConventionalDataset(s, p, pt, measured.contiguous())

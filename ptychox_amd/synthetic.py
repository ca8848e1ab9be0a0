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


def _poisson_amplitude(inten: torch.Tensor, photons: float, seed: int) -> torch.Tensor:
    scale = photons / float(inten.max())
    gen = torch.Generator(device=inten.device).manual_seed(seed)
    return torch.sqrt(torch.poisson(inten * scale, generator=gen) / scale).contiguous()


@dataclass
class FpmDataset:
    object_field: torch.Tensor   # [H, W] complex64 ground truth
    pupil: torch.Tensor          # [h, w] complex64
    leds: torch.Tensor           # [B, 2] int64 crop starts (x, y), shifted-spectrum coordinates
    measured: torch.Tensor       # [B, h, w] float32 amplitudes


@torch.no_grad()
def fpm_dataset(object_n: int, pupil_n: int, pupil_radius: float, n_side: int, pitch: int, device, seed: int = 0,
                noise: float = 0.01, sigma: float = 3.0) -> FpmDataset:
    """Fourier ptychography: smooth random object (amplitude U[0.3, 1], phase
    U[-1, 1] rad), circular pupil with a random defocus ``exp(i c (2 rho^2 - 1))``,
    ``c ~ U[-1, 1]``, an ``n_side`` x ``n_side`` LED grid on a ``pitch`` px
    crop pitch centred in the spectrum, intensities ``|ifft2(fftshift(crop q))|^2``
    plus Gaussian noise at ``noise`` x the mean."""
    rng = np.random.default_rng(seed)
    obj = smooth_complex_object(rng, object_n, sigma, amp=(0.3, 1.0), phase=(-1.0, 1.0))
    yy, xx = np.mgrid[0:pupil_n, 0:pupil_n].astype(np.float64) - pupil_n / 2
    rho2 = (xx**2 + yy**2) / pupil_radius**2
    c = rng.uniform(-1, 1)
    pupil = ((rho2 <= 1.0) * np.exp(1j * c * (2 * rho2 - 1))).astype(np.complex64)
    center = (object_n - pupil_n) // 2
    g = center + pitch * (np.arange(n_side) - n_side // 2)
    gy, gx = np.meshgrid(g, g, indexing="ij")
    leds = np.stack([gx.ravel(), gy.ravel()], axis=1).astype(np.int64)
    dev = torch.device(device)
    o = torch.from_numpy(obj).to(dev)
    q = torch.from_numpy(pupil).to(dev)
    lt = torch.from_numpy(leds).to(dev)
    spec = torch.fft.fft2(o)
    rows = torch.arange(pupil_n, device=dev)
    yy_i = (lt[:, 1, None, None] + rows[None, :, None] + object_n // 2) % object_n
    xx_i = (lt[:, 0, None, None] + rows[None, None, :] + object_n // 2) % object_n
    sub = spec.reshape(-1)[yy_i * object_n + xx_i]
    inten = torch.fft.ifft2(torch.fft.fftshift(sub * q, dim=(-2, -1))).abs().square()
    gen = torch.Generator(device=dev).manual_seed(seed)
    inten = inten + noise * inten.mean() * torch.randn(inten.shape, generator=gen, device=dev)
    return FpmDataset(o, q, lt, torch.sqrt(inten.clamp_min(0)).contiguous())


def angular_spectrum(n: int, lambda_z: float) -> np.ndarray:
    f = np.fft.fftfreq(n)
    return np.exp(1j * np.pi * lambda_z * (f[:, None] ** 2 + f[None, :] ** 2)).astype(np.complex64)


@dataclass
class NearFieldDataset:
    sample: torch.Tensor
    probe: torch.Tensor
    transfer: torch.Tensor
    positions: torch.Tensor
    measured: torch.Tensor


@torch.no_grad()
def nearfield_dataset(object_n: int, probe_n: int, n_pos: int, spread: int, fresnel: float, photons: float, device,
                      seed: int = 0) -> NearFieldDataset:
    """Near-field ptychography: smooth random object, a diffuser probe (unit
    amplitude disc, phase U[0, 2 pi) smoothed sigma = 2 px), ``n_pos`` window
    corners ``(object_n - probe_n)/2 + U{-spread..spread}``, angular-spectrum
    transfer with ``lambda z = (probe_n/2)^2 / fresnel``, Poisson noise."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    obj = smooth_complex_object(rng, object_n, 4.0)
    yy, xx = np.mgrid[0:probe_n, 0:probe_n].astype(np.float64) - probe_n / 2
    disc = (xx**2 + yy**2) <= (probe_n / 2) ** 2
    ph = _rescale(gaussian_filter(rng.uniform(0, 2 * np.pi, (probe_n, probe_n)), 2.0, mode="wrap"), 0, 2 * np.pi)
    probe = (disc * np.exp(1j * ph)).astype(np.complex64)
    c = (object_n - probe_n) // 2
    pos = (c + rng.integers(-spread, spread + 1, (n_pos, 2))).astype(np.int64)
    t = angular_spectrum(probe_n, (probe_n / 2) ** 2 / fresnel)
    dev = torch.device(device)
    s, p, tt, pt = (torch.from_numpy(a).to(dev) for a in (obj, probe, t, pos))
    rows = torch.arange(probe_n, device=dev)
    idx = (pt[:, 1, None, None] + rows[None, :, None]) * object_n + pt[:, 0, None, None] + rows[None, None, :]
    inten = torch.fft.ifft2(tt * torch.fft.fft2(s.reshape(-1)[idx] * p)).abs().square()
    return NearFieldDataset(s, p, tt, pt, _poisson_amplitude(inten, photons, seed))


@dataclass
class CodedDataset:
    object_field: torch.Tensor
    coded_surface: torch.Tensor
    transfer_1: torch.Tensor
    transfer_2: torch.Tensor
    shifts: torch.Tensor         # [B, 2] float32 normalised shifts
    downsample: int
    measured: torch.Tensor


@torch.no_grad()
def coded_dataset(n: int, n_side: int, step: int, lambda_z: float, photons: float, device, seed: int = 0,
                  downsample: int = 1) -> CodedDataset:
    """Coded ptychography: smooth random object, a binary {0,1} mask on 2 px
    cells times a U[0, pi) phase, H1 = H2 = angular spectrum (``lambda_z``),
    an ``n_side`` x ``n_side`` grid of integer translations on a ``step`` px
    pitch (``shifts = -t / n``), Poisson noise on the RMS-downsampled intensity."""
    rng = np.random.default_rng(seed)
    obj = smooth_complex_object(rng, n, 4.0)
    cells = (rng.random((n // 2, n // 2)) < 0.5).astype(np.float64)
    phase = rng.uniform(0, np.pi, (n // 2, n // 2))
    mask = (np.kron(cells, np.ones((2, 2))) * np.exp(1j * np.kron(phase, np.ones((2, 2))))).astype(np.complex64)
    t = angular_spectrum(n, lambda_z)
    g = step * (np.arange(n_side) - n_side // 2)
    gy, gx = np.meshgrid(g, g, indexing="ij")
    shifts = (-np.stack([gx.ravel(), gy.ravel()], axis=1) / n).astype(np.float32)
    dev = torch.device(device)
    o, c, tt, sh = (torch.from_numpy(a).to(dev) for a in (obj, mask, t, shifts))
    f = torch.fft.fftfreq(n, device=dev, dtype=torch.float64) * n
    yy, xx = torch.meshgrid(f, f, indexing="ij")
    spec = torch.fft.fft2(o)
    amps = []
    for k in range(sh.shape[0]):
        arg = 2 * np.pi * (sh[k, 0].double() * xx + sh[k, 1].double() * yy)
        x = torch.fft.ifft2(spec * tt * torch.polar(torch.ones_like(arg), arg).to(torch.complex64))
        z = torch.fft.ifft2(torch.fft.fft2(x * c) * tt)
        i2 = z.abs().square()
        if downsample > 1:
            i2 = torch.nn.functional.avg_pool2d(i2[None, None], downsample)[0, 0]
        amps.append(i2)
    inten = torch.stack(amps)
    return CodedDataset(o, c, tt, tt, sh, downsample, _poisson_amplitude(inten, photons, seed))

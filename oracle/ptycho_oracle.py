"""FP64 NumPy oracle for the ptychography forward/adjoint path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg, never by the shipped package.

Conventions (all from the reference torch path, file:line relative to the
reference tree):

* F = unnormalised 2-D DFT (``numpy.fft.fft2``), F^-1 = ``ifft2`` (1/N),
  S / S^-1 = ``fftshift`` / ``ifftshift`` over the last two axes.
* Complex gradients follow PyTorch: ``grad = dL/dRe z + i dL/dIm z``
  (= 2 dL/d conj(z)).  For ``y = F x`` the adjoint is ``N * ifft2``
  (``python/ptychoX/base/ops.py:95``, ``python/ptychoX/fpm/ops.py:97``).
* Positions are ``[B, 2]`` zero-based ``(x, y)``; x is the column
  (``python/ptychoX/base/ops.py:61-62``).
* The phase divisor is ``max(|z|, FLT_EPSILON)`` (``base/ops.py:93``,
  ``fpm/ops.py:86``); the coded RMS divisor is ``max(amp, 1e-7)``
  (``python/ptychoX/csrc/coded_kernel.cu:227``).

Pinning: every function here is checked in ``tests/test_oracle.py`` by at least
one of brute-force DFT / index loops, central finite differences, FP64
dot-product adjoint identities, closed-form special cases, Parseval, or the
reference's own ``python/test_ops.py`` cases.  None is "parity unpinned".
"""

from __future__ import annotations

import numpy as np

FLT_EPS = float(np.finfo(np.float32).eps)  # 1.1920929e-07, base/ops.py:93
CODED_AMP_EPS = 1.0e-7  # coded_kernel.cu:227

__all__ = [
    "FLT_EPS",
    "CODED_AMP_EPS",
    "fftshift2",
    "ifftshift2",
    "dft2_bruteforce",
    "base_crop",
    "base_forward",
    "base_backward",
    "fpm_flat_indices",
    "fpm_forward",
    "fpm_backward",
    "shift_phase_ramp",
    "coded_forward",
    "coded_backward",
    "nearfield_forward",
    "nearfield_backward",
    "MATLAB_AMP_EPS",
    "nearfield_matlab_forward",
    "nearfield_matlab_backward",
    "LOSS_KINDS",
    "stencil_loss",
    "hessian_loss",
    "mse_amplitude_loss",
    "base_train_step",
    "fpm_train_step",
    "nearfield_train_step",
    "coded_train_step",
]


# ----------------------------------------------------------------------------
# FFT helpers
# ----------------------------------------------------------------------------

def fftshift2(x: np.ndarray) -> np.ndarray:
    """``torch.fft.fftshift(x, dim=(-2, -1))`` (``base/ops.py:65``)."""
    return np.fft.fftshift(x, axes=(-2, -1))


def ifftshift2(x: np.ndarray) -> np.ndarray:
    """``torch.fft.ifftshift(x, dim=(-2, -1))`` (``base/ops.py:95``)."""
    return np.fft.ifftshift(x, axes=(-2, -1))


def dft2_bruteforce(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Explicit O(N^2) 2-D DFT over the last two axes, FP64.

    Forward: ``X[k,l] = sum_{m,n} x[m,n] exp(-2 pi i (k m / H + l n / W))``;
    inverse uses ``+`` and divides by ``H W`` (numpy/torch ``ifft2``).
    Used only to pin ``numpy.fft`` on small frames.
    """
    x = np.asarray(x, dtype=np.complex128)
    h, w = x.shape[-2:]
    sign = 1.0 if inverse else -1.0
    ky = np.arange(h)
    kx = np.arange(w)
    wy = np.exp(sign * 2j * np.pi * np.outer(ky, ky) / h)
    wx = np.exp(sign * 2j * np.pi * np.outer(kx, kx) / w)
    out = np.einsum("km,...mn,nl->...kl", wy, x, wx)
    if inverse:
        out = out / (h * w)
    return out


# ----------------------------------------------------------------------------
# Conventional ptychography (python/ptychoX/base/ops.py:42-107)
# ----------------------------------------------------------------------------

def base_crop(sample: np.ndarray, positions_xy: np.ndarray, h: int, w: int) -> np.ndarray:
    """Patch stack ``P_b = sample[y_b:y_b+h, x_b:x_b+w]`` (``base/ops.py:60-63``)."""
    positions_xy = np.asarray(positions_xy, dtype=np.int64)
    out = np.empty((positions_xy.shape[0], h, w), dtype=np.complex128)
    for b, (x, y) in enumerate(positions_xy.tolist()):
        patch = sample[y : y + h, x : x + w]
        if patch.shape != (h, w):
            raise ValueError("scan window leaves the sample")
        out[b] = patch
    return out


def base_forward(sample: np.ndarray, probe: np.ndarray, positions_xy: np.ndarray) -> dict:
    """``A_b = |S(F(P_b * p))|`` (``base/ops.py:60-66``)."""
    sample = np.asarray(sample, dtype=np.complex128)
    probe = np.asarray(probe, dtype=np.complex128)
    h, w = probe.shape
    patches = base_crop(sample, positions_xy, h, w)
    far_field = fftshift2(np.fft.fft2(patches * probe[None]))
    return {"amplitude": np.abs(far_field), "far_field": far_field, "patches": patches}


def base_backward(
    grad_output: np.ndarray,
    sample_shape: tuple[int, int],
    probe: np.ndarray,
    patches: np.ndarray,
    far_field: np.ndarray,
    positions_xy: np.ndarray,
) -> tuple[np.ndarray, np.ndarray]:
    """Hand-written backward of ``ptychoX_base`` (``base/ops.py:93-107``).

    ``Gamma = G * Phi / max(|Phi|, eps)``; ``chi = N * ifft2(ifftshift(Gamma))``;
    ``grad_sample[win_b] += chi_b * conj(p)`` and ``grad_probe += chi_b * conj(P_b)``,
    both accumulated in ascending view order like the reference loop (100-105).
    """
    probe = np.asarray(probe, dtype=np.complex128)
    h, w = probe.shape
    grad_output = np.asarray(grad_output, dtype=np.float64)
    phase = far_field / np.maximum(np.abs(far_field), FLT_EPS)
    grad_far = grad_output * phase
    grad_exit = np.fft.ifft2(ifftshift2(grad_far)) * (h * w)
    grad_sample = np.zeros(sample_shape, dtype=np.complex128)
    grad_probe = np.zeros_like(probe)
    pconj = np.conj(probe)
    for b, (x, y) in enumerate(np.asarray(positions_xy, dtype=np.int64).tolist()):
        grad_sample[y : y + h, x : x + w] += grad_exit[b] * pconj
        grad_probe += grad_exit[b] * np.conj(patches[b])
    return grad_sample, grad_probe


# ----------------------------------------------------------------------------
# Fourier ptychography (python/ptychoX/fpm/ops.py:46-124)
# ----------------------------------------------------------------------------

def fpm_flat_indices(object_shape, low_shape, led_indices_xy: np.ndarray) -> np.ndarray:
    """Wrap-around flat indices of each sub-spectrum (``fpm/ops.py:113-124``)."""
    oh, ow = object_shape
    lh, lw = low_shape
    led = np.asarray(led_indices_xy, dtype=np.int64)
    rows = np.arange(lh)
    cols = np.arange(lw)
    yy = (led[:, 1, None, None] + rows[None, :, None] + oh // 2) % oh
    xx = (led[:, 0, None, None] + cols[None, None, :] + ow // 2) % ow
    return yy * ow + xx


def fpm_forward(object_field: np.ndarray, pupil: np.ndarray, led_indices_xy: np.ndarray) -> dict:
    """``E_b = ifft2(S(sub_b * q))``, ``A_b = |E_b|`` (``fpm/ops.py:56-59``)."""
    object_field = np.asarray(object_field, dtype=np.complex128)
    pupil = np.asarray(pupil, dtype=np.complex128)
    object_fft = np.fft.fft2(object_field)
    idx = fpm_flat_indices(object_field.shape, pupil.shape, led_indices_xy)
    sub = object_fft.reshape(-1)[idx]
    image_field = np.fft.ifft2(fftshift2(sub * pupil[None]))
    return {
        "amplitude": np.abs(image_field),
        "image_field": image_field,
        "sub_spectrum": sub,
        "object_fft": object_fft,
    }


def fpm_backward(
    grad_output: np.ndarray,
    object_shape: tuple[int, int],
    pupil: np.ndarray,
    sub_spectrum: np.ndarray,
    image_field: np.ndarray,
    led_indices_xy: np.ndarray,
) -> tuple[np.ndarray, np.ndarray]:
    """Hand-written backward of ``ptychoX_fourier`` (``fpm/ops.py:85-97``)."""
    pupil = np.asarray(pupil, dtype=np.complex128)
    lh, lw = pupil.shape
    oh, ow = object_shape
    phase = image_field / np.maximum(np.abs(image_field), FLT_EPS)
    grad_image = np.asarray(grad_output, dtype=np.float64) * phase
    grad_low = ifftshift2(np.fft.fft2(grad_image) / (lh * lw))
    grad_pupil = np.sum(grad_low * np.conj(sub_spectrum), axis=0)
    idx = fpm_flat_indices(object_shape, pupil.shape, led_indices_xy).reshape(-1)
    vals = (grad_low * np.conj(pupil)[None]).reshape(-1)
    grad_fft = np.zeros(oh * ow, dtype=np.complex128)
    np.add.at(grad_fft, idx, vals)  # index_add_ (fpm/ops.py:95)
    grad_object = np.fft.ifft2(grad_fft.reshape(oh, ow)) * (oh * ow)
    return grad_object, grad_pupil


# ----------------------------------------------------------------------------
# Coded ptychography (python/ptychoX/coded/ops.py:136-171)
# ----------------------------------------------------------------------------

def shift_phase_ramp(height: int, width: int, shifts_xy: np.ndarray) -> np.ndarray:
    """``R_b = exp(i 2 pi (dx_b fx + dy_b fy))`` with signed integer frequencies
    ``f = fftfreq(n) * n`` (``coded/ops.py:157-171``)."""
    shifts = np.asarray(shifts_xy, dtype=np.float64)
    fy = np.fft.fftfreq(height) * height
    fx = np.fft.fftfreq(width) * width
    arg = 2.0 * np.pi * (shifts[:, 0, None, None] * fx[None, None, :] + shifts[:, 1, None, None] * fy[None, :, None])
    return np.exp(1j * arg)


def coded_forward(
    object_field: np.ndarray,
    coded_surface: np.ndarray,
    transfer_1: np.ndarray,
    transfer_2: np.ndarray,
    shifts_xy: np.ndarray,
    downsample: int,
) -> dict:
    """Coded forward (``coded/ops.py:136-154``).

    ``X_b = ifft2(F(O) H1 R_b)``, ``Z_b = ifft2(fft2(X_b C) H2)``,
    ``A_b = sqrt(max(avgpool_ds |Z_b|^2, 0))``.
    """
    o = np.asarray(object_field, dtype=np.complex128)
    h, w = o.shape
    ds = int(downsample)
    ramp = shift_phase_ramp(h, w, shifts_xy)
    spec = np.fft.fft2(o)
    x_fwd = np.fft.ifft2(spec[None] * np.asarray(transfer_1)[None] * ramp)
    after = x_fwd * np.asarray(coded_surface)[None]
    sensor = np.fft.ifft2(np.fft.fft2(after) * np.asarray(transfer_2)[None])
    inten = np.abs(sensor) ** 2
    b = inten.shape[0]
    pooled = inten.reshape(b, h // ds, ds, w // ds, ds).mean(axis=(2, 4))
    return {
        "amplitude": np.sqrt(np.maximum(pooled, 0.0)),
        "object_spectrum": spec,
        "x_forward": x_fwd,
        "sensor_field": sensor,
        "ramp": ramp,
    }


def coded_backward(
    grad_output: np.ndarray,
    fwd: dict,
    coded_surface: np.ndarray,
    transfer_1: np.ndarray,
    transfer_2: np.ndarray,
    downsample: int,
) -> tuple[np.ndarray, np.ndarray]:
    """Adjoint of :func:`coded_forward` w.r.t. object and coded surface.

    The reference torch path differentiates ``coded/ops.py:136-154`` by
    autograd (routing at ``coded/ops.py:45-53``); the CUDA path spells the same
    chain out (``python/ptychoX/csrc/coded_kernel.cu:347-424``):
    ``gZ = Z G_up / (ds^2 max(A_up, 1e-7))``; ``gY = ifft2(conj(H2) fft2(gZ))``;
    ``gC = sum_b gY conj(X_b)``; ``gX = gY conj(C)``;
    ``gSpec = sum_b conj(H1 R_b) fft2(gX_b) / N``; ``gO = N ifft2(gSpec)``.
    """
    ds = int(downsample)
    z = fwd["sensor_field"]
    amp = fwd["amplitude"]
    x_fwd = fwd["x_forward"]
    ramp = fwd["ramp"]
    b, h, w = z.shape
    n = h * w
    g = np.asarray(grad_output, dtype=np.float64)
    scale = g / (ds * ds * np.maximum(amp, CODED_AMP_EPS))
    scale_up = np.repeat(np.repeat(scale, ds, axis=1), ds, axis=2)
    gz = z * scale_up
    gy = np.fft.ifft2(np.conj(transfer_2)[None] * np.fft.fft2(gz))
    g_coded = np.sum(gy * np.conj(x_fwd), axis=0)
    gx = gy * np.conj(coded_surface)[None]
    g_spec = np.sum(np.conj(transfer_1)[None] * np.conj(ramp) * np.fft.fft2(gx), axis=0) / n
    g_object = np.fft.ifft2(g_spec) * n
    return g_object, g_coded


# ----------------------------------------------------------------------------
# MATLAB near-field operator (full-frame form), restated step by step from the
# MEX sources in their own operation order and normalisations
# (matlab/PtychoX_updated/@adPtychoX_nearf/adPtychoX_nearf.m:46-91,
# scripts/forward/fwd_adPtychoX_nearf.cu, scripts/backward/bwd_adPtychoX_nearf.cu,
# scripts/kernels/kernels_nearf.cuh).  Arrays in logical (row, column) order;
# x pairs with the column frequency.
# ----------------------------------------------------------------------------

MATLAB_AMP_EPS = 1.0e-3  # kernels_nearf.cuh:523 (out + 0.001f)


def _ifft_unnorm(x: np.ndarray) -> np.ndarray:
    """cuFFT's CUFFT_INVERSE: the unnormalised inverse DFT (``N * ifft2``)."""
    n = x.shape[-1] * x.shape[-2]
    return np.fft.ifft2(x) * n


def _signed_freq_phase(h: int, w: int, shifts_xy: np.ndarray) -> np.ndarray:
    """``2 pi (kx dx + ky dy)`` with signed integer frequencies
    (``i - N`` for ``2 i >= N``, kernels_nearf.cuh:57-64)."""
    ky = np.arange(h) - (2 * np.arange(h) >= h) * h
    kx = np.arange(w) - (2 * np.arange(w) >= w) * w
    sh = np.asarray(shifts_xy, dtype=np.float64)
    return 2.0 * np.pi * (sh[:, 0, None, None] * kx[None, None, :] + sh[:, 1, None, None] * ky[None, :, None])


def nearfield_matlab_forward(wavefront1, wavefront2, h_d2, shifts_xy, downsample) -> dict:
    """``observe = adPtychoX_nearf.forward(wavefront1, wavefront2, H_d2, shiftsPos)``
    (fwd_adPtychoX_nearf.cu:58-114):
    W1 = fft2(w1); X_b = ifft_unnorm(W1 e^{+i phi_b}) / N (kept as X_record);
    Z_b = ifft_unnorm(fft2(X_b w2) H2 / N); observe_b = sqrt(mean_{ds x ds} |Z_b|^2)."""
    w1 = np.asarray(wavefront1, dtype=np.complex128)
    h, w = w1.shape
    n = h * w
    ds = int(downsample)
    spec = np.fft.fft2(w1)                                                    # cu:72-77
    x = spec[None] * np.exp(1j * _signed_freq_phase(h, w, shifts_xy))        # cuh:34-80
    x = _ifft_unnorm(x)                                                       # cu:86-91
    x_record = x / n                                                          # cuh:256-291 (scaled X written back)
    z = np.fft.fft2(x_record * np.asarray(wavefront2)[None])                  # cuh:284-288, cu:101-106
    z = z * np.asarray(h_d2)[None] / n                                        # cuh:306-344
    z = _ifft_unnorm(z)                                                       # cu:112-117
    b = z.shape[0]
    pooled = (np.abs(z) ** 2).reshape(b, h // ds, ds, w // ds, ds).mean(axis=(2, 4))  # cuh:454-488
    return {"observe": np.sqrt(pooled), "x_record": x_record, "x_forward": z}


def nearfield_matlab_backward(dl_dout, fwd: dict, wavefront2, h_d2, shifts_xy, downsample):
    """``[dldw1, dldw2] = adPtychoX_nearf.backward(...)`` (bwd_adPtychoX_nearf.cu):
    P_b = Z_b (G_b / (observe_b + 0.001))_up (cuh:507-530, no 1/ds^2);
    Y_b = ifft_unnorm(fft2(P_b) conj(H2) / N) (cuh:346-392);
    dldw2 = sum_b conj(Y_b) X_b and F_b = fft2(Y_b conj(w2)) (cuh:394-436);
    S = sum_b conj(F_b e^{-i phi_b}) / N (cuh:186-240); dldw1 = fft2(S)."""
    ds = int(downsample)
    z = fwd["x_forward"]
    x_record = fwd["x_record"]
    b, h, w = z.shape
    n = h * w
    g = np.asarray(dl_dout, dtype=np.float64) / (fwd["observe"] + MATLAB_AMP_EPS)
    p = z * np.repeat(np.repeat(g, ds, axis=1), ds, axis=2)
    y = _ifft_unnorm(np.fft.fft2(p) * np.conj(np.asarray(h_d2))[None] / n)
    dldw2 = np.sum(np.conj(y) * x_record, axis=0)
    f = np.fft.fft2(y * np.conj(np.asarray(wavefront2))[None])
    s = np.sum(np.conj(f * np.exp(-1j * _signed_freq_phase(h, w, shifts_xy))), axis=0) / n
    dldw1 = np.fft.fft2(s)
    return dldw1, dldw2


# ----------------------------------------------------------------------------
# Near-field ptychography, crop form (semantics: matlab/PtychoX_updated/scripts/
# forward/fwd_adPtychoX_nearf.cu:58-114 -- FFT, transfer multiply, inverse FFT,
# detector magnitude -- applied to a probe-illuminated crop as the conventional
# operator does)
# ----------------------------------------------------------------------------

def nearfield_forward(sample, probe, transfer, positions_xy) -> dict:
    """``Psi_b = ifft2(T * fft2(P_b * p))``, ``A_b = |Psi_b|`` (T has DC at [0, 0])."""
    sample = np.asarray(sample, dtype=np.complex128)
    probe = np.asarray(probe, dtype=np.complex128)
    h, w = probe.shape
    patches = base_crop(sample, positions_xy, h, w)
    psi = np.fft.ifft2(np.asarray(transfer)[None] * np.fft.fft2(patches * probe[None]))
    return {"amplitude": np.abs(psi), "field": psi, "patches": patches}


def nearfield_backward(grad_output, sample_shape, probe, transfer, patches, field, positions_xy):
    """Adjoint: ``chi_b = ifft2(conj(T) fft2(G Psi/max(|Psi|,eps)))`` (no factor N,
    because the adjoint of ``F^-1 D_T F`` is ``F^-1 D_conj(T) F``); object and
    probe accumulation as in :func:`base_backward`."""
    probe = np.asarray(probe, dtype=np.complex128)
    h, w = probe.shape
    phase = field / np.maximum(np.abs(field), FLT_EPS)
    gam = np.asarray(grad_output, dtype=np.float64) * phase
    chi = np.fft.ifft2(np.conj(transfer)[None] * np.fft.fft2(gam))
    grad_sample = np.zeros(sample_shape, dtype=np.complex128)
    grad_probe = np.zeros_like(probe)
    pconj = np.conj(probe)
    for b, (x, y) in enumerate(np.asarray(positions_xy, dtype=np.int64).tolist()):
        grad_sample[y : y + h, x : x + w] += chi[b] * pconj
        grad_probe += chi[b] * np.conj(patches[b])
    return grad_sample, grad_probe


# ----------------------------------------------------------------------------
# Amplitude loss of the training drivers (python/train_ptychography.py:555,
# python/train_fpm.py:261, python/train_coded.py:235: F.mse_loss, mean)
# ----------------------------------------------------------------------------

# ----------------------------------------------------------------------------
# Alternative losses (python/train_coded.py:101-107; matlab/Examples/
# Ptychography/+loss_fun/{l2,fd,tv,sparse,hessian}_loss.m)
# ----------------------------------------------------------------------------

LOSS_KINDS = ("l2", "fd_iso", "fd_aniso", "grad_l2", "sparse_iso", "sparse_aniso")


def _sgn(z):
    m = np.abs(z)
    return np.where(m > 0, z / np.where(m > 0, m, 1.0), 0.0)


def stencil_loss(kind: str, d: np.ndarray, eps: float = 0.0):
    """Per-frame sums of the penalty of ``d`` (real or complex, [..., h, w])
    and its gradient (PyTorch convention), with forward differences
    ``dx = roll(d, -1, -1) - d``, ``dy = roll(d, -1, -2) - d``
    (train_coded.py:102-104, +loss_fun/tv_loss.m:2-3)."""
    d = np.asarray(d)
    if kind == "l2":
        return np.sum(np.abs(d) ** 2, axis=(-2, -1)), 2.0 * d
    dx = np.roll(d, -1, axis=-1) - d
    dy = np.roll(d, -1, axis=-2) - d
    mx, my = np.abs(dx), np.abs(dy)
    if kind in ("fd_iso", "sparse_iso"):
        s = np.sqrt(mx ** 2 + my ** 2 + eps)
        k = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
        phi = s
        if kind == "sparse_iso":
            e = np.exp(-8.0 * s)
            phi, k = 1.0 - e, 8.0 * e * k
        ax, ay = dx * k, dy * k
    elif kind in ("fd_aniso", "sparse_aniso"):
        phi = mx + my
        ax, ay = _sgn(dx), _sgn(dy)
        if kind == "sparse_aniso":
            e = np.exp(-8.0 * phi)
            phi, ax, ay = 1.0 - e, 8.0 * e * ax, 8.0 * e * ay
    elif kind == "grad_l2":
        phi = mx ** 2 + my ** 2
        ax, ay = 2.0 * dx, 2.0 * dy
    else:
        raise ValueError(kind)
    grad = np.roll(ax, 1, axis=-1) + np.roll(ay, 1, axis=-2) - ax - ay
    return np.sum(phi, axis=(-2, -1)), grad


def hessian_loss(dw: np.ndarray, kind: str) -> float:
    """+loss_fun/hessian_loss.m: second differences dwdxx, dwdyy, dwdxy of the
    wrapped first differences; 'isotropic' sqrt(|dxx|^2 + |dyy|^2 + 2|dxy|^2),
    'anisotropic' |dxx| + |dyy| + 2|dxy|, 'L2' |dx|^2 + |dy|^2; summed."""
    dw = np.asarray(dw)
    dx = np.roll(dw, -1, axis=-1) - dw   # dW(:,[2:end,1]) - dW
    dy = np.roll(dw, -1, axis=-2) - dw   # dW([2:end,1],:) - dW
    if kind == "L2":
        return float(np.sum(np.abs(dx) ** 2 + np.abs(dy) ** 2))
    dxx = np.roll(dx, -1, axis=-1) - dx
    dyy = np.roll(dy, -1, axis=-2) - dy
    dxy = np.roll(dx, -1, axis=-2) - dx
    if kind == "isotropic":
        return float(np.sum(np.sqrt(np.abs(dxx) ** 2 + np.abs(dyy) ** 2 + 2 * np.abs(dxy) ** 2)))
    if kind == "anisotropic":
        return float(np.sum(np.abs(dxx) + np.abs(dyy) + 2 * np.abs(dxy)))
    raise ValueError(kind)


def mse_amplitude_loss(amplitude: np.ndarray, measured: np.ndarray):
    """Returns ``(L, L_b, G)`` with ``L = mean (A - Y)^2``, per-position
    ``L_b = mean_i (A_b - Y_b)^2`` and ``G = dL/dA = 2 (A - Y) / numel``."""
    a = np.asarray(amplitude, dtype=np.float64)
    y = np.asarray(measured, dtype=np.float64)
    d = a - y
    per = np.mean(d.reshape(d.shape[0], -1) ** 2, axis=1)
    return float(np.mean(d**2)), per, 2.0 * d / d.size


def base_train_step(sample, probe, positions_xy, measured):
    """Loss + gradients of one conventional mini-batch step
    (``train_ptychography.py:549-558``: op -> mse_loss -> backward)."""
    fwd = base_forward(sample, probe, positions_xy)
    loss, per, g = mse_amplitude_loss(fwd["amplitude"], measured)
    gs, gp = base_backward(g, np.shape(sample), probe, fwd["patches"], fwd["far_field"], positions_xy)
    return loss, per, gs, gp


def fpm_train_step(object_field, pupil, led_indices_xy, measured):
    """Loss + gradients of one FPM step (``train_fpm.py:257-263``)."""
    fwd = fpm_forward(object_field, pupil, led_indices_xy)
    loss, per, g = mse_amplitude_loss(fwd["amplitude"], measured)
    go, gp = fpm_backward(g, np.shape(object_field), pupil, fwd["sub_spectrum"], fwd["image_field"], led_indices_xy)
    return loss, per, go, gp


def nearfield_train_step(sample, probe, transfer, positions_xy, measured):
    """Loss + gradients of one near-field mini-batch step."""
    fwd = nearfield_forward(sample, probe, transfer, positions_xy)
    loss, per, g = mse_amplitude_loss(fwd["amplitude"], measured)
    gs, gp = nearfield_backward(g, np.shape(sample), probe, transfer, fwd["patches"], fwd["field"], positions_xy)
    return loss, per, gs, gp


def coded_train_step(object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample, measured):
    """Loss + gradients of one coded step (``train_coded.py:231-238``)."""
    fwd = coded_forward(object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample)
    loss, per, g = mse_amplitude_loss(fwd["amplitude"], measured)
    go, gc = coded_backward(g, fwd, coded_surface, transfer_1, transfer_2, downsample)
    return loss, per, go, gc

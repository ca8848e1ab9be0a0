"""Near-field ptychography operator ``ptychoX_nearfield``.

The reference ships near-field only as a MATLAB/MEX operator
(``matlab/PtychoX_updated/@adPtychoX_nearf``); this is its Python operator on
the same surface as the others: each probe-illuminated scan window is
propagated by the angular-spectrum transfer function ``transfer`` (DC at
``[0, 0]``) and the detector records the amplitude,
``A_b = |ifft2(transfer * fft2(crop_b(sample) * probe))|``.
"""

from __future__ import annotations

import math

import torch

from . import _lib
from ._utils import (
    begin_window_check,
    can_use_native,
    check_complex_scan_inputs,
    check_windows,
    end_window_check,
    should_recompute,
    scatter_flags,
    require_native_frame,
)
from .base import crop_indices


def ptychoX_nearfield(
    sample: torch.Tensor,
    probe: torch.Tensor,
    transfer: torch.Tensor,
    positions_xy: torch.Tensor,
    use_cuda_kernels: bool = False,
) -> torch.Tensor:
    """Near-field amplitudes ``[batch, h, w]``.

    Args:
        sample: complex object ``[H, W]``.
        probe: complex probe ``[h, w]``.
        transfer: complex propagation transfer function ``[h, w]``, DC at ``[0, 0]``
            (see :func:`angular_spectrum_transfer`).
        positions_xy: zero-based ``(x, y)`` window corners ``[batch, 2]``.
        use_cuda_kernels: run the native HIP kernels (complex64 GPU tensors).
    """
    return PtychoXNearFieldFunction.apply(sample, probe, transfer, positions_xy, use_cuda_kernels)


nearfield_ptychography = ptychoX_nearfield


def clear_cuda_plan_cache() -> None:
    """Kept for symmetry with the other operators; nothing is cached."""


def angular_spectrum_transfer(h: int, w: int, lambda_z: float, device=None, dtype=torch.complex64) -> torch.Tensor:
    """``exp(i pi lambda_z (fx^2 + fy^2))`` on the unshifted ``fftfreq`` grid, in
    pixel units (``lambda_z`` in px^2; Fresnel number ``a^2 / lambda_z``)."""
    fy = torch.fft.fftfreq(h, device=device, dtype=torch.float64)
    fx = torch.fft.fftfreq(w, device=device, dtype=torch.float64)
    arg = math.pi * lambda_z * (fy[:, None] ** 2 + fx[None, :] ** 2)
    return torch.polar(torch.ones_like(arg), arg).to(dtype)


def _check_transfer(transfer: torch.Tensor, probe: torch.Tensor) -> None:
    if transfer.shape != probe.shape:
        raise ValueError("transfer must have the same shape as probe.")
    if not transfer.is_complex():
        raise TypeError("transfer must be a complex tensor.")
    if transfer.device != probe.device:
        raise ValueError("transfer must live on the same device as probe.")


# ----------------------------------------------------------------------------- torch path

def torch_forward(sample, probe, transfer, positions_xy):
    h, w = probe.shape
    patches = sample.reshape(-1)[crop_indices(positions_xy, h, w, sample.shape[1])]
    field = torch.fft.ifft2(transfer.unsqueeze(0) * torch.fft.fft2(patches * probe.unsqueeze(0)))
    return patches, field, field.abs()


def torch_backward(grad_output, patches, probe, transfer, field, positions_xy, sample_shape):
    h, w = probe.shape
    eps = torch.finfo(field.real.dtype).eps
    gamma = grad_output * field / field.abs().clamp_min(eps)
    chi = torch.fft.ifft2(transfer.conj().unsqueeze(0) * torch.fft.fft2(gamma))
    idx = crop_indices(positions_xy, h, w, sample_shape[1])
    gs = torch.zeros(sample_shape[0] * sample_shape[1], dtype=probe.dtype, device=probe.device)
    gs.index_add_(0, idx.reshape(-1), (chi * probe.conj().unsqueeze(0)).reshape(-1))
    return gs.reshape(sample_shape), torch.sum(chi * patches.conj(), dim=0)


# ----------------------------------------------------------------------------- native path

def native_forward(sample, probe, transfer, positions_xy):
    H, W = sample.shape
    h, w = probe.shape
    B = positions_xy.shape[0]
    dev = sample.device
    field = torch.empty((B, h, w), dtype=torch.complex64, device=dev)
    amplitude = torch.empty((B, h, w), dtype=torch.float32, device=dev)
    _lib.call(
        "ptx_nearfield_forward",
        _lib.ptr(sample), H, W, _lib.ptr(probe), h, w, _lib.ptr(transfer), _lib.ptr(positions_xy), B,
        _lib.ptr(field), _lib.ptr(amplitude), _lib.stream_of(dev),
    )
    return field, amplitude


def native_backward(grad_output, field, sample, probe, transfer, positions_xy, flags=None):
    H, W = sample.shape
    h, w = probe.shape
    B = positions_xy.shape[0]
    dev = sample.device
    if flags is None:
        flags = scatter_flags()
    ws = _lib.workspace(_lib.query("ptx_nearfield_backward_workspace", H, W, h, w, B, flags), dev)
    gs = torch.empty((H, W), dtype=torch.complex64, device=dev)
    gp = torch.empty((h, w), dtype=torch.complex64, device=dev)
    _lib.call(
        "ptx_nearfield_backward",
        _lib.ptr(grad_output), _lib.ptr(field), _lib.ptr(sample), H, W, _lib.ptr(probe), h, w, _lib.ptr(transfer),
        _lib.ptr(positions_xy), B, _lib.ptr(gs), _lib.ptr(gp), _lib.ptr(ws), ws.numel(), flags, _lib.stream_of(dev),
    )
    return gs, gp


class PtychoXNearFieldFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sample, probe, transfer, positions_xy, use_cuda_kernels=False):  # type: ignore[override]
        check_complex_scan_inputs(sample, probe, positions_xy)
        _check_transfer(transfer, probe)
        h, w = probe.shape
        native = can_use_native(sample, probe, use_cuda_kernels)
        # checked before the move to the device: CPU positions cost no
        # synchronisation; GPU positions on the HIP path are checked beside
        # the forward kernels (end_window_check below)
        pending = begin_window_check(positions_xy, h, w, *sample.shape) if native else None
        if not native:
            check_windows(positions_xy, h, w, *sample.shape)
        positions_xy = positions_xy.to(device=sample.device, dtype=torch.long).contiguous()
        sample = sample.contiguous()
        probe = probe.contiguous()
        transfer = transfer.to(probe.dtype).contiguous()
        ctx.sample_shape = tuple(sample.shape)
        ctx.recompute = False
        if native:
            require_native_frame(h, w, "probe")
            field, amplitude = native_forward(sample, probe, transfer, positions_xy)
            end_window_check(pending)
            ctx.recompute = should_recompute(field.numel() * 8, sample.device)
            ctx.save_for_backward(sample, probe, transfer, None if ctx.recompute else field, positions_xy)
            ctx.native = True
        else:
            patches, field, amplitude = torch_forward(sample, probe, transfer, positions_xy)
            ctx.save_for_backward(patches, probe, transfer, field, positions_xy)
            ctx.native = False
        return amplitude

    @staticmethod
    def backward(ctx, grad_output):  # type: ignore[override]
        first, probe, transfer, field, positions_xy = ctx.saved_tensors
        grad_output = grad_output.to(dtype=torch.float32 if ctx.native else field.real.dtype).contiguous()
        if ctx.native:
            if ctx.recompute:
                field, _ = native_forward(first, probe, transfer, positions_xy)
            gs, gp = native_backward(grad_output, field, first, probe, transfer, positions_xy)
        else:
            gs, gp = torch_backward(grad_output, first, probe, transfer, field, positions_xy, ctx.sample_shape)
        return gs, gp, None, None, None


class NearFieldStep:
    """Fused near-field training step on the HIP path (``ptx_nearfield_step``)."""

    def __init__(self, sample_shape, probe_shape, batch, device, scatter=None, keep_amplitude=False):
        self.H, self.W = (int(s) for s in sample_shape)
        self.h, self.w = (int(s) for s in probe_shape)
        require_native_frame(self.h, self.w, "probe")
        self.batch = int(batch)
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.flags = scatter_flags(scatter)
        self.ws = _lib.workspace(self._ws_bytes(), self.device)
        self.loss = torch.empty(self.batch, dtype=torch.float32, device=self.device)
        self.amplitude = (
            torch.empty((self.batch, self.h, self.w), dtype=torch.float32, device=self.device)
            if keep_amplitude
            else None
        )

    def _ws_bytes(self) -> int:
        # the size depends on the FFT schedule in the tuning table, so it is
        # re-queried (host-side, no GPU work) before every call
        return _lib.query("ptx_nearfield_step_workspace", self.H, self.W, self.h, self.w, self.batch, self.flags)

    def __call__(self, sample, probe, transfer, positions_xy, measured, grad_sample, grad_probe, loss_scale=None,
                 batch=None, accumulate=False):
        B = self.batch if batch is None else int(batch)
        if B > self.batch:
            raise ValueError(f"batch {B} exceeds the planned batch {self.batch}")
        for name, t, dt in (("sample", sample, torch.complex64), ("probe", probe, torch.complex64),
                            ("transfer", transfer, torch.complex64), ("positions_xy", positions_xy, torch.int64),
                            ("measured", measured, torch.float32), ("grad_sample", grad_sample, torch.complex64),
                            ("grad_probe", grad_probe, torch.complex64)):
            if t.device != self.device or t.dtype != dt or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {dt} tensor on {self.device}")
        if loss_scale is None:
            loss_scale = 1.0 / (B * self.h * self.w) if B > 0 else 0.0
        flags = self.flags | (_lib.PTX_FLAG_ACCUMULATE if accumulate else 0)
        if self._ws_bytes() > self.ws.numel():
            self.ws = _lib.workspace(self._ws_bytes(), self.device)
        _lib.call(
            "ptx_nearfield_step",
            _lib.ptr(sample), self.H, self.W, _lib.ptr(probe), self.h, self.w, _lib.ptr(transfer),
            _lib.ptr(positions_xy), B, _lib.ptr(measured), float(loss_scale), _lib.ptr(self.loss),
            _lib.ptr(self.amplitude), _lib.ptr(grad_sample), _lib.ptr(grad_probe), _lib.ptr(self.ws),
            self.ws.numel(), flags, _lib.stream_of(self.device),
        )
        return self.loss[:B]

"""Conventional ptychography operator ``ptychoX_base`` (drop-in for
``python/ptychoX/base/ops.py``).

``use_cuda_kernels=True`` on complex64 GPU tensors runs the hand-written HIP
path (``ptx_base_forward`` / ``ptx_base_backward`` in libptychox_hip.so);
otherwise the same math runs as unfused torch ops (the reference's torch
fallback, vectorised: no per-view Python loop and no device-to-host sync).
"""

from __future__ import annotations

import torch

from . import _lib
from ._utils import (
    begin_window_check,
    can_use_native,
    check_complex_scan_inputs,
    check_windows,
    end_window_check,
    require_native_frame,
    scatter_flags,
    should_recompute,
)


def ptychoX_base(
    sample: torch.Tensor,
    probe: torch.Tensor,
    positions_xy: torch.Tensor,
    use_cuda_kernels: bool = False,
) -> torch.Tensor:
    """Return ``abs(fftshift(fft2(crop(sample) * probe)))``
    (``python/ptychoX/base/ops.py:14-28``).

    Args:
        sample: complex object tensor, shape ``[sample_h, sample_w]``.
        probe: complex probe tensor, shape ``[probe_h, probe_w]``.
        positions_xy: zero-based ``(x, y)`` scan positions, shape ``[batch, 2]``.
        use_cuda_kernels: use the native HIP kernels (complex64 GPU tensors).
    """
    return PtychoXBaseFunction.apply(sample, probe, positions_xy, use_cuda_kernels)


ptychography_base = ptychoX_base


def clear_cuda_plan_cache() -> None:
    """Kept for drop-in compatibility (``base/ops.py:34-37``): the HIP FFTs are
    compile-time plans, so there is no cache to clear."""


# ----------------------------------------------------------------------------- torch path

def crop_indices(positions_xy: torch.Tensor, h: int, w: int, W: int) -> torch.Tensor:
    """Flat indices ``[B, h, w]`` of the scan windows in a row-major sample of width W."""
    rows = torch.arange(h, device=positions_xy.device)
    cols = torch.arange(w, device=positions_xy.device)
    y = positions_xy[:, 1, None, None] + rows[None, :, None]
    x = positions_xy[:, 0, None, None] + cols[None, None, :]
    return y * W + x


def torch_forward(sample, probe, positions_xy):
    h, w = probe.shape
    idx = crop_indices(positions_xy, h, w, sample.shape[1])
    patches = sample.reshape(-1)[idx]
    far_field = torch.fft.fftshift(torch.fft.fft2(patches * probe.unsqueeze(0)), dim=(-2, -1))
    return patches, far_field, far_field.abs()


def torch_backward(grad_output, patches, probe, far_field, positions_xy, sample_shape):
    h, w = probe.shape
    eps = torch.finfo(far_field.real.dtype).eps
    phase = far_field / far_field.abs().clamp_min(eps)
    grad_exit = torch.fft.ifft2(torch.fft.ifftshift(grad_output * phase, dim=(-2, -1))) * (h * w)
    idx = crop_indices(positions_xy, h, w, sample_shape[1])
    grad_sample = torch.zeros(sample_shape[0] * sample_shape[1], dtype=probe.dtype, device=probe.device)
    grad_sample.index_add_(0, idx.reshape(-1), (grad_exit * probe.conj().unsqueeze(0)).reshape(-1))
    grad_probe = torch.sum(grad_exit * patches.conj(), dim=0)
    return grad_sample.reshape(sample_shape), grad_probe


# ----------------------------------------------------------------------------- native path

def native_forward(sample, probe, positions_xy):
    """``(far_field, amplitude)`` from ``ptx_base_forward``."""
    H, W = sample.shape
    h, w = probe.shape
    B = positions_xy.shape[0]
    far_field = torch.empty((B, h, w), dtype=torch.complex64, device=sample.device)
    amplitude = torch.empty((B, h, w), dtype=torch.float32, device=sample.device)
    _lib.call(
        "ptx_base_forward",
        _lib.ptr(sample), H, W, _lib.ptr(probe), h, w, _lib.ptr(positions_xy), B,
        _lib.ptr(far_field), _lib.ptr(amplitude), _lib.stream_of(sample.device),
    )
    return far_field, amplitude


def native_backward(grad_output, far_field, sample, probe, positions_xy, flags=None):
    H, W = sample.shape
    h, w = probe.shape
    B = positions_xy.shape[0]
    if flags is None:
        flags = scatter_flags()
    dev = sample.device
    nbytes = _lib.query("ptx_base_backward_workspace", H, W, h, w, B, flags)
    ws = _lib.workspace(nbytes, dev)
    grad_sample = torch.empty((H, W), dtype=torch.complex64, device=dev)
    grad_probe = torch.empty((h, w), dtype=torch.complex64, device=dev)
    _lib.call(
        "ptx_base_backward",
        _lib.ptr(grad_output), _lib.ptr(far_field), _lib.ptr(sample), H, W, _lib.ptr(probe), h, w,
        _lib.ptr(positions_xy), B, _lib.ptr(grad_sample), _lib.ptr(grad_probe),
        _lib.ptr(ws), ws.numel(), flags, _lib.stream_of(dev),
    )
    return grad_sample, grad_probe


class PtychoXBaseFunction(torch.autograd.Function):
    """Forward/backward of ``ptychoX_base`` (``python/ptychoX/base/ops.py:40-107``)."""

    @staticmethod
    def forward(ctx, sample, probe, positions_xy, use_cuda_kernels=False):  # type: ignore[override]
        check_complex_scan_inputs(sample, probe, positions_xy)
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
        ctx.sample_shape = tuple(sample.shape)
        ctx.recompute = False
        if native:
            require_native_frame(h, w, "probe")
            far_field, amplitude = native_forward(sample, probe, positions_xy)
            end_window_check(pending)
            # the far field is the one per-pattern tensor kept for the backward
            # (the patches are re-gathered from the sample); recomputed instead
            # when it does not fit (python/ptychoX/base/ops.py:69 saves both)
            ctx.recompute = should_recompute(far_field.numel() * 8, sample.device)
            ctx.save_for_backward(sample, probe, None if ctx.recompute else far_field, positions_xy)
            ctx.native = True
        else:
            patches, far_field, amplitude = torch_forward(sample, probe, positions_xy)
            ctx.save_for_backward(patches, probe, far_field, positions_xy)
            ctx.native = False
        return amplitude

    @staticmethod
    def backward(ctx, grad_output):  # type: ignore[override]
        first, probe, far_field, positions_xy = ctx.saved_tensors
        grad_output = grad_output.to(dtype=torch.float32 if ctx.native else far_field.real.dtype).contiguous()
        if ctx.native:
            if ctx.recompute:
                far_field, _ = native_forward(first, probe, positions_xy)
            grad_sample, grad_probe = native_backward(grad_output, far_field, first, probe, positions_xy)
        else:
            grad_sample, grad_probe = torch_backward(
                grad_output, first, probe, far_field, positions_xy, ctx.sample_shape
            )
        return grad_sample, grad_probe, None, None


# ----------------------------------------------------------------------------- fused step

class BaseStep:
    """Fused conventional training step on the HIP path (``ptx_base_step``):
    forward, amplitude MSE loss against ``measured`` and both gradients in three
    passes, with a reusable workspace (no allocation per call).

    ``loss_scale`` defaults to ``1/(B*h*w)``, i.e. ``F.mse_loss(..., "mean")`` of
    the reference drivers (python/train_ptychography.py:555).
    """

    def __init__(self, sample_shape, probe_shape, batch, device, scatter=None, keep_amplitude=False,
                 path=None):
        self.H, self.W = (int(s) for s in sample_shape)
        self.h, self.w = (int(s) for s in probe_shape)
        require_native_frame(self.h, self.w, "probe")
        self.batch = int(batch)
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        # path: None (the library's choice: whole-frame kernel for 32^2 probes,
        # row / column passes otherwise), "three_pass" or "whole_frame"
        paths = {None: 0, "three_pass": _lib.PTX_FLAG_THREE_PASS, "whole_frame": _lib.PTX_FLAG_WHOLE_FRAME}
        if path not in paths:
            raise ValueError(f"path must be one of {list(paths)}, got {path!r}")
        self.flags = scatter_flags(scatter) | paths[path]
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
        return _lib.query("ptx_base_step_workspace", self.H, self.W, self.h, self.w, self.batch, self.flags)

    def __call__(self, sample, probe, positions_xy, measured, grad_sample, grad_probe, loss_scale=None,
                 batch=None, accumulate=False):
        B = self.batch if batch is None else int(batch)
        for name, t, dt in (("sample", sample, torch.complex64), ("probe", probe, torch.complex64),
                            ("positions_xy", positions_xy, torch.int64), ("measured", measured, torch.float32),
                            ("grad_sample", grad_sample, torch.complex64), ("grad_probe", grad_probe, torch.complex64)):
            if t.device != self.device or t.dtype != dt or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {dt} tensor on {self.device}")
        if B > self.batch:
            raise ValueError(f"batch {B} exceeds the planned batch {self.batch}")
        if loss_scale is None:
            loss_scale = 1.0 / (B * self.h * self.w) if B > 0 else 0.0
        flags = self.flags | (_lib.PTX_FLAG_ACCUMULATE if accumulate else 0)
        if self._ws_bytes() > self.ws.numel():
            self.ws = _lib.workspace(self._ws_bytes(), self.device)
        _lib.call(
            "ptx_base_step",
            _lib.ptr(sample), self.H, self.W, _lib.ptr(probe), self.h, self.w, _lib.ptr(positions_xy), B,
            _lib.ptr(measured), float(loss_scale), _lib.ptr(self.loss), _lib.ptr(self.amplitude),
            _lib.ptr(grad_sample), _lib.ptr(grad_probe), _lib.ptr(self.ws), self.ws.numel(), flags,
            _lib.stream_of(self.device),
        )
        return self.loss[:B]

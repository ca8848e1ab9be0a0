"""Coded ptychography operator ``ptychoX_coded`` (drop-in for
``python/ptychoX/coded/ops.py``).

As in the reference (``coded/ops.py:45-53``), the torch path returns the forward
of plain torch ops and lets autograd differentiate it; ``use_cuda_kernels=True``
on complex64 GPU tensors runs the HIP forward/adjoint behind an
``autograd.Function``.
"""

from __future__ import annotations

import math

import torch
import torch.nn.functional as F

from . import _lib
from ._utils import can_use_native, require_native_frame, should_recompute


def ptychoX_coded(
    object_field: torch.Tensor,
    coded_surface: torch.Tensor,
    transfer_1: torch.Tensor,
    transfer_2: torch.Tensor,
    shifts_xy: torch.Tensor,
    downsample: int,
    use_cuda_kernels: bool = False,
) -> torch.Tensor:
    """Forward model for coded ptychography (``python/ptychoX/coded/ops.py:22-62``).

    Args:
        object_field: high-resolution complex object, ``[H, W]``.
        coded_surface: high-resolution complex coded surface, ``[H, W]``.
        transfer_1: propagation transfer function before the coded surface.
        transfer_2: propagation transfer function after the coded surface.
        shifts_xy: normalized subpixel shifts, ``[batch, 2]``, ``(dx, dy)``.
        downsample: integer detector downsample factor.

    Returns:
        RMS downsampled amplitudes, shape ``[batch, H/downsample, W/downsample]``.
    """
    check_inputs(object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample)
    if not can_use_native(object_field, coded_surface, use_cuda_kernels):
        return torch_forward(object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample)[3]
    return PtychoXCodedFunction.apply(
        object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample, use_cuda_kernels
    )


coded_ptychography = ptychoX_coded


def clear_cuda_plan_cache() -> None:
    """Kept for drop-in compatibility (``coded/ops.py:16-19``); nothing is cached."""


def check_inputs(object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample) -> None:
    """Same checks and messages as ``python/ptychoX/coded/ops.py:174-200``."""
    if object_field.ndim != 2:
        raise ValueError("object_field must be a 2D complex tensor.")
    for name, tensor in (("coded_surface", coded_surface), ("transfer_1", transfer_1), ("transfer_2", transfer_2)):
        if tensor.shape != object_field.shape:
            raise ValueError(f"{name} must have the same shape as object_field.")
        if tensor.device != object_field.device:
            raise ValueError(f"{name} must live on the same device as object_field.")
    if not object_field.is_complex() or not coded_surface.is_complex():
        raise TypeError("object_field and coded_surface must be complex tensors.")
    if not transfer_1.is_complex() or not transfer_2.is_complex():
        raise TypeError("transfer_1 and transfer_2 must be complex tensors.")
    if shifts_xy.ndim != 2 or shifts_xy.shape[-1] != 2:
        raise ValueError("shifts_xy must have shape [batch, 2].")
    if object_field.shape[0] % downsample != 0 or object_field.shape[1] % downsample != 0:
        raise ValueError("object_field shape must be divisible by downsample.")


def shift_phase_ramp(height, width, shifts_xy, device, dtype) -> torch.Tensor:
    """``exp(i 2 pi (dx fx + dy fy))`` with signed integer frequencies
    (``coded/ops.py:157-171``)."""
    fy = torch.fft.fftfreq(height, d=1.0, device=device).to(dtype) * height
    fx = torch.fft.fftfreq(width, d=1.0, device=device).to(dtype) * width
    yy, xx = torch.meshgrid(fy, fx, indexing="ij")
    arg = 2.0 * math.pi * (shifts_xy[:, 0, None, None] * xx.unsqueeze(0) + shifts_xy[:, 1, None, None] * yy.unsqueeze(0))
    return torch.exp(1j * arg)


def torch_forward(object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample):
    """Unfused torch ops (``coded/ops.py:136-154``); returns
    ``(object_spectrum, x_forward, sensor_field, amplitude)``."""
    shifts_xy = shifts_xy.to(device=object_field.device, dtype=object_field.real.dtype)
    h, w = object_field.shape
    ramp = shift_phase_ramp(h, w, shifts_xy, object_field.device, object_field.real.dtype)
    spec = torch.fft.fft2(object_field)
    x_fwd = torch.fft.ifft2(spec.unsqueeze(0) * transfer_1.unsqueeze(0) * ramp)
    sensor = torch.fft.ifft2(torch.fft.fft2(x_fwd * coded_surface.unsqueeze(0)) * transfer_2.unsqueeze(0))
    pooled = F.avg_pool2d(sensor.abs().square().unsqueeze(1), kernel_size=downsample, stride=downsample)
    return spec, x_fwd, sensor, torch.sqrt(pooled[:, 0].clamp_min(0.0))


def native_forward(object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample):
    H, W = object_field.shape
    B = shifts_xy.shape[0]
    ds = int(downsample)
    dev = object_field.device
    spec = torch.empty((H, W), dtype=torch.complex64, device=dev)
    x_fwd = torch.empty((B, H, W), dtype=torch.complex64, device=dev)
    sensor = torch.empty((B, H, W), dtype=torch.complex64, device=dev)
    amp = torch.empty((B, H // ds, W // ds), dtype=torch.float32, device=dev)
    ws = _lib.workspace(_lib.query("ptx_coded_workspace", H, W, B, ds, 0), dev)
    _lib.call(
        "ptx_coded_forward",
        _lib.ptr(object_field), H, W, _lib.ptr(coded_surface), _lib.ptr(transfer_1), _lib.ptr(transfer_2),
        _lib.ptr(shifts_xy), B, ds, _lib.ptr(spec), _lib.ptr(x_fwd), _lib.ptr(sensor), _lib.ptr(amp),
        _lib.ptr(ws), ws.numel(), _lib.stream_of(dev),
    )
    return spec, x_fwd, sensor, amp


def native_backward(grad_output, spec, x_fwd, sensor, coded_surface, transfer_1, transfer_2, shifts_xy, downsample,
                    flags=0):
    H, W = spec.shape
    B = shifts_xy.shape[0]
    ds = int(downsample)
    dev = spec.device
    ws = _lib.workspace(_lib.query("ptx_coded_workspace", H, W, B, ds, 0), dev)
    grad_object = torch.empty((H, W), dtype=torch.complex64, device=dev)
    grad_coded = torch.empty((H, W), dtype=torch.complex64, device=dev)
    _lib.call(
        "ptx_coded_backward",
        _lib.ptr(grad_output), _lib.ptr(spec), _lib.ptr(x_fwd), _lib.ptr(sensor), _lib.ptr(coded_surface),
        _lib.ptr(transfer_1), _lib.ptr(transfer_2), _lib.ptr(shifts_xy), B, H, W, ds, _lib.ptr(grad_object),
        _lib.ptr(grad_coded), _lib.ptr(ws), ws.numel(), int(flags), _lib.stream_of(dev),
    )
    return grad_object, grad_coded


class PtychoXCodedFunction(torch.autograd.Function):
    """Native forward/backward of ``ptychoX_coded`` (``coded/ops.py:68-133``)."""

    @staticmethod
    def forward(ctx, object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample,
                use_cuda_kernels=False):  # type: ignore[override]
        check_inputs(object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample)
        require_native_frame(*object_field.shape, "object_field")
        object_field = object_field.contiguous()
        coded_surface = coded_surface.contiguous()
        transfer_1 = transfer_1.to(torch.complex64).contiguous()
        transfer_2 = transfer_2.to(torch.complex64).contiguous()
        shifts_xy = shifts_xy.to(device=object_field.device, dtype=torch.float32).contiguous()
        spec, x_fwd, sensor, amp = native_forward(
            object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample
        )
        # x_forward and sensor_field are the per-translation tensors the
        # reference keeps (python/ptychoX/coded/ops.py:108); recomputed from the
        # inputs instead when they do not fit
        ctx.recompute = should_recompute((x_fwd.numel() + sensor.numel()) * 8, object_field.device)
        if ctx.recompute:
            ctx.save_for_backward(object_field, None, None, coded_surface, transfer_1, transfer_2, shifts_xy)
        else:
            ctx.save_for_backward(spec, x_fwd, sensor, coded_surface, transfer_1, transfer_2, shifts_xy)
        ctx.downsample = int(downsample)
        return amp

    @staticmethod
    def backward(ctx, grad_output):  # type: ignore[override]
        spec, x_fwd, sensor, coded_surface, t1, t2, shifts_xy = ctx.saved_tensors
        grad_output = grad_output.to(torch.float32).contiguous()
        if ctx.recompute:  # spec holds the object itself
            spec, x_fwd, sensor, _ = native_forward(spec, coded_surface, t1, t2, shifts_xy, ctx.downsample)
        go, gc = native_backward(grad_output, spec, x_fwd, sensor, coded_surface, t1, t2, shifts_xy, ctx.downsample)
        return go, gc, None, None, None, None, None


class CodedStep:
    """Fused coded training step on the HIP path (``ptx_coded_step``).

    ``loss`` is the coded driver's ``--loss`` (python/train_coded.py:234-237):
    ``"mse"`` (``F.mse_loss``, fused into the step, ``loss_scale`` defaults to
    ``1/(B h w)``) or ``"fd-isotropic"`` / ``"fd-anisotropic"`` (``fd_loss``,
    a sum, ``loss_scale`` defaults to 1), which run the native forward, the
    stencil loss writing the upstream gradient, and the native adjoint."""

    LOSSES = ("mse", "fd-isotropic", "fd-anisotropic")

    def __init__(self, shape, batch, downsample, device, keep_amplitude=False, loss="mse"):
        self.H, self.W = (int(s) for s in shape)
        require_native_frame(self.H, self.W, "object_field")
        self.batch = int(batch)
        self.ds = int(downsample)
        if loss not in self.LOSSES:
            raise ValueError(f"loss must be one of {self.LOSSES}")
        self.loss_name = loss
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.ws = _lib.workspace(self._ws_bytes(), self.device)
        self.loss = torch.empty(self.batch, dtype=torch.float32, device=self.device)
        h, w = self.H // self.ds, self.W // self.ds
        self.amplitude = (
            torch.empty((self.batch, h, w), dtype=torch.float32, device=self.device)
            if keep_amplitude or loss != "mse"
            else None
        )
        if loss != "mse":
            dev, B = self.device, self.batch
            self._spec = torch.empty((self.H, self.W), dtype=torch.complex64, device=dev)
            self._x_fwd = torch.empty((B, self.H, self.W), dtype=torch.complex64, device=dev)
            self._sensor = torch.empty_like(self._x_fwd)
            self._g = torch.empty((B, h, w), dtype=torch.float32, device=dev)
            self._loss_ws = _lib.workspace(_lib.query("ptx_loss_workspace", B, h, w), dev)

    def _ws_bytes(self) -> int:
        # the size depends on the FFT schedule in the tuning table, so it is
        # re-queried (host-side, no GPU work) before every call
        return _lib.query("ptx_coded_workspace", self.H, self.W, self.batch, self.ds, 1 if self.loss_name == "mse" else 0)

    def __call__(self, object_field, coded_surface, transfer_1, transfer_2, shifts_xy, measured, grad_object,
                 grad_coded, loss_scale=None, batch=None, accumulate=False):
        B = self.batch if batch is None else int(batch)
        if B > self.batch:
            raise ValueError(f"batch {B} exceeds the planned batch {self.batch}")
        for name, t, dt in (("object_field", object_field, torch.complex64),
                            ("coded_surface", coded_surface, torch.complex64),
                            ("transfer_2", transfer_2, torch.complex64),
                            ("shifts_xy", shifts_xy, torch.float32), ("measured", measured, torch.float32),
                            ("grad_object", grad_object, torch.complex64),
                            ("grad_coded", grad_coded, torch.complex64)):
            if t.device != self.device or t.dtype != dt or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {dt} tensor on {self.device}")
        if transfer_1 is not None and (transfer_1.device != self.device or transfer_1.dtype != torch.complex64
                                       or not transfer_1.is_contiguous()):
            raise ValueError(f"transfer_1 must be None or a contiguous complex64 tensor on {self.device}")
        h, w = self.H // self.ds, self.W // self.ds
        flags = _lib.PTX_FLAG_ACCUMULATE if accumulate else 0
        if self._ws_bytes() > self.ws.numel():
            self.ws = _lib.workspace(self._ws_bytes(), self.device)
        stream = _lib.stream_of(self.device)
        if self.loss_name == "mse":
            if loss_scale is None:
                loss_scale = 1.0 / (B * h * w) if B > 0 else 0.0
            _lib.call(
                "ptx_coded_step",
                _lib.ptr(object_field), self.H, self.W, _lib.ptr(coded_surface), _lib.ptr(transfer_1),
                _lib.ptr(transfer_2), _lib.ptr(shifts_xy), B, self.ds, _lib.ptr(measured), float(loss_scale),
                _lib.ptr(self.loss), _lib.ptr(self.amplitude), _lib.ptr(grad_object), _lib.ptr(grad_coded),
                _lib.ptr(self.ws), self.ws.numel(), flags, stream,
            )
            return self.loss[:B]
        from .losses import FD_EPS, native_amplitude_loss

        if loss_scale is None:
            loss_scale = 1.0
        iso = self.loss_name == "fd-isotropic"
        _lib.call(
            "ptx_coded_forward",
            _lib.ptr(object_field), self.H, self.W, _lib.ptr(coded_surface), _lib.ptr(transfer_1),
            _lib.ptr(transfer_2), _lib.ptr(shifts_xy), B, self.ds, _lib.ptr(self._spec), _lib.ptr(self._x_fwd),
            _lib.ptr(self._sensor), _lib.ptr(self.amplitude), _lib.ptr(self.ws), self.ws.numel(), stream,
        )
        native_amplitude_loss("fd_iso" if iso else "fd_aniso", self.amplitude[:B], measured[:B], loss_scale,
                              FD_EPS if iso else 0.0, self.loss[:B], self._g[:B], self._loss_ws)
        _lib.call(
            "ptx_coded_backward",
            _lib.ptr(self._g), _lib.ptr(self._spec), _lib.ptr(self._x_fwd), _lib.ptr(self._sensor),
            _lib.ptr(coded_surface), _lib.ptr(transfer_1), _lib.ptr(transfer_2), _lib.ptr(shifts_xy), B, self.H,
            self.W, self.ds, _lib.ptr(grad_object), _lib.ptr(grad_coded), _lib.ptr(self.ws), self.ws.numel(), flags,
            stream,
        )
        return self.loss[:B]

"""Fourier ptychography operator ``ptychoX_fourier`` (drop-in for
``python/ptychoX/fpm/ops.py``).

``use_cuda_kernels=True`` on complex64 GPU tensors runs the HIP path
(``ptx_fpm_forward`` / ``ptx_fpm_backward``); otherwise the same math runs as
unfused torch ops.
"""

from __future__ import annotations

import torch

from . import _lib
from ._utils import can_use_native, check_complex_scan_inputs, require_native_frame, scatter_flags, should_recompute


def ptychoX_fourier(
    object_field: torch.Tensor,
    pupil: torch.Tensor,
    led_indices_xy: torch.Tensor,
    use_cuda_kernels: bool = False,
) -> torch.Tensor:
    """Fourier ptychography amplitude operator (``python/ptychoX/fpm/ops.py:14-25``).

    The PyTorch convention is row-major and zero-based:
    ``led_indices_xy[:, 0]`` is x and ``led_indices_xy[:, 1]`` is y (the crop
    start in fftshifted-spectrum coordinates; indices wrap around).
    """
    return PtychoXFPMFunction.apply(object_field, pupil, led_indices_xy, use_cuda_kernels)


fourier_ptychography = ptychoX_fourier


def clear_cuda_plan_cache() -> None:
    """Kept for drop-in compatibility (``fpm/ops.py:31-34``); nothing is cached."""


# ----------------------------------------------------------------------------- torch path

def fpm_flat_indices(object_shape, low_shape, led_indices_xy: torch.Tensor) -> torch.Tensor:
    """Wrap-around flat indices of the sub-spectra (``fpm/ops.py:113-124``)."""
    oh, ow = object_shape
    lh, lw = low_shape
    rows = torch.arange(lh, device=led_indices_xy.device, dtype=torch.long)
    cols = torch.arange(lw, device=led_indices_xy.device, dtype=torch.long)
    yy = (led_indices_xy[:, 1, None, None] + rows[None, :, None] + oh // 2) % oh
    xx = (led_indices_xy[:, 0, None, None] + cols[None, None, :] + ow // 2) % ow
    return yy * ow + xx


def torch_forward(object_field, pupil, led):
    object_fft = torch.fft.fft2(object_field)
    idx = fpm_flat_indices(object_field.shape, pupil.shape, led)
    sub = object_fft.reshape(-1)[idx]
    image_field = torch.fft.ifft2(torch.fft.fftshift(sub * pupil.unsqueeze(0), dim=(-2, -1)))
    return sub, image_field, image_field.abs()


def torch_backward(grad_output, sub, pupil, image_field, led, object_shape):
    lh, lw = pupil.shape
    oh, ow = object_shape
    eps = torch.finfo(image_field.real.dtype).eps
    phase = image_field / image_field.abs().clamp_min(eps)
    grad_low = torch.fft.ifftshift(torch.fft.fft2(grad_output * phase) / (lh * lw), dim=(-2, -1))
    grad_pupil = torch.sum(grad_low * sub.conj(), dim=0)
    idx = fpm_flat_indices(object_shape, pupil.shape, led)
    flat = torch.zeros(oh * ow, dtype=pupil.dtype, device=pupil.device)
    flat.index_add_(0, idx.reshape(-1), (grad_low * pupil.conj().unsqueeze(0)).reshape(-1))
    grad_object = torch.fft.ifft2(flat.reshape(oh, ow)) * (oh * ow)
    return grad_object, grad_pupil


# ----------------------------------------------------------------------------- native path

def native_forward(object_field, pupil, led):
    H, W = object_field.shape
    h, w = pupil.shape
    B = led.shape[0]
    dev = object_field.device
    object_fft = torch.empty((H, W), dtype=torch.complex64, device=dev)
    image_field = torch.empty((B, h, w), dtype=torch.complex64, device=dev)
    amplitude = torch.empty((B, h, w), dtype=torch.float32, device=dev)
    _lib.call(
        "ptx_fpm_forward",
        _lib.ptr(object_field), H, W, _lib.ptr(pupil), h, w, _lib.ptr(led), B,
        _lib.ptr(object_fft), _lib.ptr(image_field), _lib.ptr(amplitude), _lib.stream_of(dev),
    )
    return object_fft, image_field, amplitude


def native_backward(grad_output, image_field, object_fft, pupil, led, flags=None):
    H, W = object_fft.shape
    h, w = pupil.shape
    B = led.shape[0]
    dev = pupil.device
    if flags is None:
        flags = scatter_flags()
    ws = _lib.workspace(_lib.query("ptx_fpm_backward_workspace", H, W, h, w, B, flags), dev)
    grad_object = torch.empty((H, W), dtype=torch.complex64, device=dev)
    grad_pupil = torch.empty((h, w), dtype=torch.complex64, device=dev)
    _lib.call(
        "ptx_fpm_backward",
        _lib.ptr(grad_output), _lib.ptr(image_field), _lib.ptr(object_fft), H, W, _lib.ptr(pupil), h, w,
        _lib.ptr(led), B, _lib.ptr(grad_object), _lib.ptr(grad_pupil), _lib.ptr(ws), ws.numel(), flags,
        _lib.stream_of(dev),
    )
    return grad_object, grad_pupil


class PtychoXFPMFunction(torch.autograd.Function):
    """Forward/backward of ``ptychoX_fourier`` (``python/ptychoX/fpm/ops.py:37-99``)."""

    @staticmethod
    def forward(ctx, object_field, pupil, led_indices_xy, use_cuda_kernels=False):  # type: ignore[override]
        check_complex_scan_inputs(object_field, pupil, led_indices_xy, "object_field", "pupil")
        led = led_indices_xy.to(device=object_field.device, dtype=torch.long).contiguous()
        object_field = object_field.contiguous()
        pupil = pupil.contiguous()
        ctx.object_shape = tuple(object_field.shape)
        if can_use_native(object_field, pupil, use_cuda_kernels):
            require_native_frame(*object_field.shape, "object_field")
            require_native_frame(*pupil.shape, "pupil")
            if pupil.shape[0] > object_field.shape[0] or pupil.shape[1] > object_field.shape[1]:
                raise ValueError("pupil must not be larger than object_field")
            object_fft, image_field, amplitude = native_forward(object_field, pupil, led)
            # the image fields are the per-LED tensor kept for the backward
            # (python/ptychoX/fpm/ops.py:62 also keeps the sub-spectra, here
            # re-gathered); recomputed from the object instead when they do not fit
            ctx.recompute = should_recompute(image_field.numel() * 8, object_field.device)
            if ctx.recompute:
                ctx.save_for_backward(object_field, pupil, None, led)
            else:
                ctx.save_for_backward(object_fft, pupil, image_field, led)
            ctx.native = True
        else:
            sub, image_field, amplitude = torch_forward(object_field, pupil, led)
            ctx.save_for_backward(sub, pupil, image_field, led)
            ctx.native = False
        return amplitude

    @staticmethod
    def backward(ctx, grad_output):  # type: ignore[override]
        first, pupil, image_field, led = ctx.saved_tensors
        grad_output = grad_output.to(dtype=torch.float32 if ctx.native else image_field.real.dtype).contiguous()
        if ctx.native:
            if ctx.recompute:  # first is the object itself
                first, image_field, _ = native_forward(first, pupil, led)
            grad_object, grad_pupil = native_backward(grad_output, image_field, first, pupil, led)
        else:
            grad_object, grad_pupil = torch_backward(grad_output, first, pupil, image_field, led, ctx.object_shape)
        return grad_object, grad_pupil, None, None


class FpmStep:
    """Fused FPM training step on the HIP path (``ptx_fpm_step``): object FFT,
    forward, amplitude MSE loss and both gradients, reusable workspace."""

    def __init__(self, object_shape, pupil_shape, batch, device, scatter=None, keep_amplitude=False):
        self.H, self.W = (int(s) for s in object_shape)
        self.h, self.w = (int(s) for s in pupil_shape)
        require_native_frame(self.H, self.W, "object_field")
        require_native_frame(self.h, self.w, "pupil")
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
        return _lib.query("ptx_fpm_step_workspace", self.H, self.W, self.h, self.w, self.batch, self.flags)

    def __call__(self, object_field, pupil, led, measured, grad_object, grad_pupil, loss_scale=None, batch=None,
                 accumulate=False):
        B = self.batch if batch is None else int(batch)
        if B > self.batch:
            raise ValueError(f"batch {B} exceeds the planned batch {self.batch}")
        for name, t, dt in (("object_field", object_field, torch.complex64), ("pupil", pupil, torch.complex64),
                            ("led_indices_xy", led, torch.int64), ("measured", measured, torch.float32),
                            ("grad_object", grad_object, torch.complex64), ("grad_pupil", grad_pupil, torch.complex64)):
            if t.device != self.device or t.dtype != dt or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {dt} tensor on {self.device}")
        if loss_scale is None:
            loss_scale = 1.0 / (B * self.h * self.w) if B > 0 else 0.0
        flags = self.flags | (_lib.PTX_FLAG_ACCUMULATE if accumulate else 0)
        if self._ws_bytes() > self.ws.numel():
            self.ws = _lib.workspace(self._ws_bytes(), self.device)
        _lib.call(
            "ptx_fpm_step",
            _lib.ptr(object_field), self.H, self.W, _lib.ptr(pupil), self.h, self.w, _lib.ptr(led), B,
            _lib.ptr(measured), float(loss_scale), _lib.ptr(self.loss), _lib.ptr(self.amplitude),
            _lib.ptr(grad_object), _lib.ptr(grad_pupil), _lib.ptr(self.ws), self.ws.numel(), flags,
            _lib.stream_of(self.device),
        )
        return self.loss[:B]

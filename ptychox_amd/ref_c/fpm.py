"""``ptychoX.fpm._C`` on this library (python/ptychoX/csrc/fpm_extension.cpp:21-65)."""

from __future__ import annotations

import torch

from . import check, invoke, ptr, scratch, stream
from .._lib import query


def forward(object: torch.Tensor, pupil: torch.Tensor, led_indices_xy: torch.Tensor) -> list[torch.Tensor]:  # noqa: A002
    """``[sub_spectrum, image_field, amplitude]`` (fpm_extension.cpp:21-32).

    The sub-spectra are not stored: slot 0 is the object spectrum
    ``fft2(object)``, from which :func:`backward` re-gathers them."""
    check(object, "object", "complex64")
    check(pupil, "pupil", "complex64")
    check(led_indices_xy, "led_indices_xy", "int64")
    obj, pupil, led = object.contiguous(), pupil.contiguous(), led_indices_xy.contiguous()
    (H, W), (h, w), B, dev = obj.shape, pupil.shape, led.shape[0], obj.device
    spectrum = torch.empty((H, W), dtype=torch.complex64, device=dev)
    image = torch.empty((B, h, w), dtype=torch.complex64, device=dev)
    amp = torch.empty((B, h, w), dtype=torch.float32, device=dev)
    invoke("ptx_fpm_forward", ptr(obj), H, W, ptr(pupil), h, w, ptr(led), B, ptr(spectrum), ptr(image), ptr(amp),
           stream(dev))
    return [spectrum, image, amp]


def backward(grad_output: torch.Tensor, sub_spectrum: torch.Tensor, pupil: torch.Tensor, image_field: torch.Tensor,
             led_indices_xy: torch.Tensor, object_h: int, object_w: int) -> list[torch.Tensor]:
    """``[grad_object, grad_pupil]`` (fpm_extension.cpp:34-60); ``sub_spectrum``
    is slot 0 of :func:`forward`'s result (the object spectrum)."""
    check(grad_output, "grad_output", "float32")
    check(sub_spectrum, "sub_spectrum", "complex64")
    check(pupil, "pupil", "complex64")
    check(image_field, "image_field", "complex64")
    check(led_indices_xy, "led_indices_xy", "int64")
    if tuple(sub_spectrum.shape) != (object_h, object_w):
        raise RuntimeError("sub_spectrum must be the object spectrum returned by forward "
                           "(this library re-gathers sub-spectra)")
    g, spec, pupil = grad_output.contiguous(), sub_spectrum.contiguous(), pupil.contiguous()
    image, led = image_field.contiguous(), led_indices_xy.contiguous()
    (h, w), B, dev = pupil.shape, led.shape[0], spec.device
    ws = scratch(query("ptx_fpm_backward_workspace", object_h, object_w, h, w, B, 0), dev)
    grad_object = torch.empty((object_h, object_w), dtype=torch.complex64, device=dev)
    grad_pupil = torch.empty((h, w), dtype=torch.complex64, device=dev)
    invoke("ptx_fpm_backward", ptr(g), ptr(image), ptr(spec), object_h, object_w, ptr(pupil), h, w, ptr(led), B,
           ptr(grad_object), ptr(grad_pupil), ptr(ws), ws.numel(), 0, stream(dev))
    return [grad_object, grad_pupil]


def clear_plan_cache() -> None:
    """fpm_extension.cpp:64: the FFT plans are compile-time here, nothing is cached."""

"""``ptychoX.coded._C`` on this library (python/ptychoX/csrc/coded_extension.cpp:26-93)."""

from __future__ import annotations

import torch

from . import check, invoke, ptr, scratch, stream
from .._lib import query


def forward(object: torch.Tensor, coded_surface: torch.Tensor, transfer_1: torch.Tensor,  # noqa: A002
            transfer_2: torch.Tensor, shifts_xy: torch.Tensor, downsample: int) -> list[torch.Tensor]:
    """``[object_spectrum, x_forward, sensor_field, amplitude]`` (coded_extension.cpp:26-56)."""
    check(object, "object", "complex64")
    check(coded_surface, "coded_surface", "complex64")
    check(transfer_1, "transfer_1", "complex64")
    check(transfer_2, "transfer_2", "complex64")
    check(shifts_xy, "shifts_xy", "float32")
    obj, cs = object.contiguous(), coded_surface.contiguous()
    t1, t2, sh = transfer_1.contiguous(), transfer_2.contiguous(), shifts_xy.contiguous()
    (H, W), B, ds, dev = obj.shape, sh.shape[0], int(downsample), obj.device
    if ds <= 0 or H % ds or W % ds:
        raise RuntimeError("object shape must be divisible by downsample")
    spec = torch.empty((H, W), dtype=torch.complex64, device=dev)
    x_fwd = torch.empty((B, H, W), dtype=torch.complex64, device=dev)
    sensor = torch.empty((B, H, W), dtype=torch.complex64, device=dev)
    amp = torch.empty((B, H // ds, W // ds), dtype=torch.float32, device=dev)
    ws = scratch(query("ptx_coded_workspace", H, W, B, ds, 0), dev)
    invoke("ptx_coded_forward", ptr(obj), H, W, ptr(cs), ptr(t1), ptr(t2), ptr(sh), B, ds, ptr(spec), ptr(x_fwd),
           ptr(sensor), ptr(amp), ptr(ws), ws.numel(), stream(dev))
    return [spec, x_fwd, sensor, amp]


def backward(grad_output: torch.Tensor, object_spectrum: torch.Tensor, x_forward: torch.Tensor,
             sensor_field: torch.Tensor, coded_surface: torch.Tensor, transfer_1: torch.Tensor,
             transfer_2: torch.Tensor, shifts_xy: torch.Tensor, downsample: int) -> list[torch.Tensor]:
    """``[grad_object, grad_coded]`` (coded_extension.cpp:58-88)."""
    check(grad_output, "grad_output", "float32")
    check(object_spectrum, "object_spectrum", "complex64")
    check(x_forward, "x_forward", "complex64")
    check(sensor_field, "sensor_field", "complex64")
    check(coded_surface, "coded_surface", "complex64")
    check(transfer_1, "transfer_1", "complex64")
    check(transfer_2, "transfer_2", "complex64")
    check(shifts_xy, "shifts_xy", "float32")
    g, spec, x_fwd = grad_output.contiguous(), object_spectrum.contiguous(), x_forward.contiguous()
    sensor, cs = sensor_field.contiguous(), coded_surface.contiguous()
    t1, t2, sh = transfer_1.contiguous(), transfer_2.contiguous(), shifts_xy.contiguous()
    (H, W), B, ds, dev = spec.shape, sh.shape[0], int(downsample), spec.device
    ws = scratch(query("ptx_coded_workspace", H, W, B, ds, 0), dev)
    grad_object = torch.empty((H, W), dtype=torch.complex64, device=dev)
    grad_coded = torch.empty((H, W), dtype=torch.complex64, device=dev)
    invoke("ptx_coded_backward", ptr(g), ptr(spec), ptr(x_fwd), ptr(sensor), ptr(cs), ptr(t1), ptr(t2), ptr(sh),
           B, H, W, ds, ptr(grad_object), ptr(grad_coded), ptr(ws), ws.numel(), 0, stream(dev))
    return [grad_object, grad_coded]


def clear_plan_cache() -> None:
    """coded_extension.cpp:92: the FFT plans are compile-time here, nothing is cached."""

"""``ptychoX.base._C`` on this library (python/ptychoX/csrc/base_extension.cpp:21-65)."""

from __future__ import annotations

import torch

from . import check, invoke, ptr, scratch, stream
from .._lib import query


def forward(sample: torch.Tensor, probe: torch.Tensor, positions_xy: torch.Tensor) -> list[torch.Tensor]:
    """``[patch_stack, far_field_shifted, amplitude]`` (base_extension.cpp:21-32).

    The patches are not stored: slot 0 is the (contiguous) sample, from which
    :func:`backward` re-gathers them."""
    check(sample, "sample", "complex64")
    check(probe, "probe", "complex64")
    check(positions_xy, "positions_xy", "int64")
    sample, probe, positions_xy = sample.contiguous(), probe.contiguous(), positions_xy.contiguous()
    (H, W), (h, w), B = sample.shape, probe.shape, positions_xy.shape[0]
    far = torch.empty((B, h, w), dtype=torch.complex64, device=sample.device)
    amp = torch.empty((B, h, w), dtype=torch.float32, device=sample.device)
    invoke("ptx_base_forward", ptr(sample), H, W, ptr(probe), h, w, ptr(positions_xy), B, ptr(far), ptr(amp),
           stream(sample.device))
    return [sample, far, amp]


def backward(grad_output: torch.Tensor, patch_stack: torch.Tensor, probe: torch.Tensor,
             far_field_shifted: torch.Tensor, positions_xy: torch.Tensor, sample_h: int,
             sample_w: int) -> list[torch.Tensor]:
    """``[grad_sample, grad_probe]`` (base_extension.cpp:34-60); ``patch_stack``
    is slot 0 of :func:`forward`'s result (the sample)."""
    check(grad_output, "grad_output", "float32")
    check(patch_stack, "patch_stack", "complex64")
    check(probe, "probe", "complex64")
    check(far_field_shifted, "far_field_shifted", "complex64")
    check(positions_xy, "positions_xy", "int64")
    if tuple(patch_stack.shape) != (sample_h, sample_w):
        raise RuntimeError("patch_stack must be the sample returned by forward (this library re-gathers patches)")
    sample, probe = patch_stack.contiguous(), probe.contiguous()
    g, far, pos = grad_output.contiguous(), far_field_shifted.contiguous(), positions_xy.contiguous()
    (h, w), B, dev = probe.shape, pos.shape[0], sample.device
    ws = scratch(query("ptx_base_backward_workspace", sample_h, sample_w, h, w, B, 0), dev)
    grad_sample = torch.empty((sample_h, sample_w), dtype=torch.complex64, device=dev)
    grad_probe = torch.empty((h, w), dtype=torch.complex64, device=dev)
    invoke("ptx_base_backward", ptr(g), ptr(far), ptr(sample), sample_h, sample_w, ptr(probe), h, w, ptr(pos), B,
           ptr(grad_sample), ptr(grad_probe), ptr(ws), ws.numel(), 0, stream(dev))
    return [grad_sample, grad_probe]


def clear_plan_cache() -> None:
    """base_extension.cpp:64: the FFT plans are compile-time here, nothing is cached."""

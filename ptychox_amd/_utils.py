"""Input validation and small helpers shared by the operators.

Messages and exception types follow the reference
(``python/ptychoX/_utils.py:24-46,99-101``) so callers see the same errors.
"""

from __future__ import annotations

import os

import torch

PTX_MIN_N = 8
PTX_MAX_N = 2048

_SCATTER_MODES = ("segments", "atomic")
_SCATTER = os.environ.get("PTYCHOX_SCATTER", "segments")
if _SCATTER not in _SCATTER_MODES:
    raise ValueError(f"PTYCHOX_SCATTER must be one of {_SCATTER_MODES}, got {_SCATTER!r}")


def check_complex_scan_inputs(
    first: torch.Tensor,
    second: torch.Tensor,
    positions_xy: torch.Tensor,
    first_name: str = "sample",
    second_name: str = "probe",
) -> None:
    """Same checks and messages as ``python/ptychoX/_utils.py:24-42``."""
    if first.ndim != 2 or second.ndim != 2:
        raise ValueError(f"{first_name} and {second_name} must be 2D complex tensors.")
    if not first.is_complex() or not second.is_complex():
        raise TypeError(f"{first_name} and {second_name} must be complex tensors.")
    if first.device != second.device:
        raise ValueError(f"{first_name} and {second_name} must live on the same device.")
    if first.dtype != second.dtype:
        raise ValueError(f"{first_name} and {second_name} must have the same dtype.")
    if positions_xy.ndim != 2 or positions_xy.shape[-1] != 2:
        raise ValueError("positions_xy must have shape [batch, 2].")
    if positions_xy.device.type != "cpu" and positions_xy.device != first.device:
        raise ValueError(f"positions_xy must be on CPU or on the same device as {first_name}.")


def can_use_native(first: torch.Tensor, second: torch.Tensor, requested: bool) -> bool:
    """``use_cuda_kernels=True`` on complex64 GPU tensors selects the HIP path
    (the rule of ``python/ptychoX/_utils.py:45-46``; the flag keeps its name
    for drop-in compatibility and means "native kernels" on ROCm)."""
    return bool(requested) and first.is_cuda and first.dtype == torch.complex64 and second.dtype == torch.complex64


def native_size_ok(n: int) -> bool:
    return PTX_MIN_N <= n <= PTX_MAX_N and (n & (n - 1)) == 0


def require_native_frame(h: int, w: int, what: str) -> None:
    if not (native_size_ok(h) and native_size_ok(w)):
        raise ValueError(
            f"native HIP path: {what} {h}x{w} unsupported; sides must be powers of two in "
            f"[{PTX_MIN_N}, {PTX_MAX_N}] (call with use_cuda_kernels=False for other sizes)"
        )


def check_windows(positions_xy: torch.Tensor, h: int, w: int, H: int, W: int) -> None:
    """Every scan window ``[y, y+h) x [x, x+w)`` must lie inside the ``H x W`` sample.

    The reference has no such check (its torch slice shrinks and ``torch.stack``
    fails; its CUDA kernel reads out of bounds, python/ptychoX/csrc/base_kernel.cu:108-137).
    One 4-number device-to-host copy when the positions live on the GPU; none
    when they live on the CPU, and no check at all while the current stream is
    being captured into a graph.
    """
    if positions_xy.numel() == 0:
        return
    if positions_xy.is_cuda and torch.cuda.is_current_stream_capturing():
        # no host synchronisation inside a stream capture: the library's own
        # contract applies (an out-of-range window gives a zero frame and no
        # gradient, without touching memory outside the sample)
        return
    lo, hi = torch.aminmax(positions_xy, dim=0)  # one reduction for both columns
    (xmin, ymin), (xmax, ymax) = torch.stack((lo, hi)).tolist()
    _raise_if_outside(xmin, ymin, xmax, ymax, h, w, H, W)


def _raise_if_outside(xmin, ymin, xmax, ymax, h, w, H, W) -> None:
    if xmin < 0 or ymin < 0 or xmax + w > W or ymax + h > H:
        raise ValueError(
            f"scan window out of range: positions x in [{xmin}, {xmax}], y in [{ymin}, {ymax}] with a "
            f"{h}x{w} probe do not fit a {H}x{W} sample."
        )


_CHECK_STREAMS: dict[int, torch.cuda.Stream] = {}


def begin_window_check(positions_xy: torch.Tensor, h: int, w: int, H: int, W: int):
    """``check_windows`` in two halves, for a caller that launches the HIP
    kernels in between (they give an out-of-range window a zero frame and
    touch no memory outside the sample, so launching them first is safe).

    Positions on the GPU: the min/max reduction runs on a side stream (after
    the work already queued on the current stream) and its four numbers are
    copied to page-locked memory; ``end_window_check`` waits for them and
    raises the same error as ``check_windows``.  The host therefore queues
    the kernels before it waits, and the device runs them meanwhile.  CPU
    positions are checked here at once (no synchronisation); an empty batch
    or a stream capture skips the check.  Returns the handle for
    ``end_window_check`` (None when nothing is pending)."""
    if positions_xy.numel() == 0:
        return None
    if not positions_xy.is_cuda:
        check_windows(positions_xy, h, w, H, W)
        return None
    if torch.cuda.is_current_stream_capturing():
        return None
    dev = positions_xy.device
    side = _CHECK_STREAMS.get(dev.index)
    if side is None:
        side = _CHECK_STREAMS.setdefault(dev.index, torch.cuda.Stream(device=dev))
    side.wait_stream(torch.cuda.current_stream(dev))
    host = torch.empty((2, 2), dtype=positions_xy.dtype, pin_memory=True)
    with torch.cuda.stream(side):
        lo, hi = torch.aminmax(positions_xy, dim=0)
        host.copy_(torch.stack((lo, hi)), non_blocking=True)
        done = torch.cuda.Event()
        done.record(side)
    return host, done, (h, w, H, W)


def end_window_check(pending) -> None:
    """Second half of ``begin_window_check``: wait for the four numbers and
    raise if a scan window leaves the sample."""
    if pending is None:
        return
    host, done, (h, w, H, W) = pending
    done.synchronize()
    (xmin, ymin), (xmax, ymax) = host.tolist()
    _raise_if_outside(xmin, ymin, xmax, ymax, h, w, H, W)


def set_scatter_mode(mode: str | None) -> None:
    """How object / spectrum gradients are accumulated on the HIP path:
    ``"segments"`` (default; sorted-segment gather in ascending view order,
    bitwise reproducible) or ``"atomic"`` (float atomic adds).  ``None``
    restores the default.  ``torch.use_deterministic_algorithms(True)``
    always selects ``"segments"``."""
    global _SCATTER
    mode = "segments" if mode is None else mode
    if mode not in _SCATTER_MODES:
        raise ValueError(f"scatter mode must be one of {_SCATTER_MODES}, got {mode!r}")
    _SCATTER = mode


def scatter_mode() -> str:
    return "segments" if torch.are_deterministic_algorithms_enabled() else _SCATTER


def scatter_flags(mode: str | None = None) -> int:
    """C-ABI flag bits for ``mode`` (default: the current :func:`scatter_mode`)."""
    from ._lib import PTX_FLAG_ATOMIC_SCATTER, PTX_FLAG_DETERMINISTIC

    if torch.are_deterministic_algorithms_enabled():
        return PTX_FLAG_DETERMINISTIC
    mode = scatter_mode() if mode is None else mode
    if mode not in _SCATTER_MODES:
        raise ValueError(f"scatter mode must be one of {_SCATTER_MODES}, got {mode!r}")
    return PTX_FLAG_ATOMIC_SCATTER if mode == "atomic" else 0


def matlab_positions_to_torch(positions_xy_1based: torch.Tensor) -> torch.Tensor:
    """Convert MATLAB one-based ``(x, y)`` scan positions to zero-based int64
    (``python/ptychoX/_utils.py:99-101``)."""
    return positions_xy_1based.to(torch.long) - 1


# ----------------------------------------------------------------------------- recompute instead of store

_RECOMPUTE: str | int = "auto"


def set_recompute(mode: str | int | None) -> None:
    """What the native autograd operators keep for their backward.

    The reference saves each call's per-pattern FFT outputs (far field, image
    field, near-field field, coded ``x_forward`` and ``sensor_field``:
    python/ptychoX/base/ops.py:69, fpm/ops.py:62, coded/ops.py:108).  Here they
    are either saved or recomputed from the saved inputs in the backward
    (bitwise the same values: the kernels are deterministic).  ``mode``:

    * ``"never"``: always save them (the reference's behaviour);
    * ``"always"``: always recompute;
    * an int ``n``: recompute when they would take more than ``n`` bytes;
    * ``"auto"`` (default, also for ``None``): recompute when they would take
      more than half of the device memory free at the forward call.
    """
    global _RECOMPUTE
    if mode is None:
        mode = "auto"
    if isinstance(mode, bool) or not (isinstance(mode, int) or mode in ("never", "always", "auto")):
        raise ValueError(f"recompute mode must be 'never', 'always', 'auto' or a byte budget, got {mode!r}")
    if isinstance(mode, int) and mode < 0:
        raise ValueError("recompute byte budget must be >= 0")
    _RECOMPUTE = mode


def recompute_mode() -> str | int:
    return _RECOMPUTE


def should_recompute(saved_bytes: int, device: torch.device, free_bytes: int | None = None) -> bool:
    """True when a forward that would save ``saved_bytes`` of per-pattern
    outputs should keep only its inputs (see :func:`set_recompute`)."""
    m = recompute_mode()
    if m == "never":
        return False
    if m == "always":
        return True
    if isinstance(m, int):
        return saved_bytes > m
    if free_bytes is None:
        if device.type != "cuda" or torch.cuda.is_current_stream_capturing():
            return False  # no device query inside a graph capture: keep the outputs
        free_bytes, _total = torch.cuda.mem_get_info(device)
    return saved_bytes > free_bytes // 2


_env = os.environ.get("PTYCHOX_RECOMPUTE")
if _env:
    set_recompute(int(_env) if _env.isdigit() else _env)

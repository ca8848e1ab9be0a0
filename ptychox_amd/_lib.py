"""ctypes binding of the C ABI in ``include/ptx.h`` (libptychox_hip.so).

The library is loaded after ``torch`` so that it binds to the HIP runtime torch
already mapped (same ``libamdhip64.so.7`` soname): one runtime, one set of
streams.  A missing or unloadable library raises -- there is no fallback.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime first)

LIB_PATH = Path(__file__).resolve().parent / "libptychox_hip.so"

PTX_FLAG_DETERMINISTIC = 1
PTX_FLAG_ACCUMULATE = 2
PTX_FLAG_ATOMIC_SCATTER = 4
PTX_FLAG_MATLAB_NEARF = 8
PTX_FLAG_THREE_PASS = 16
PTX_FLAG_WHOLE_FRAME = 32

PTX_LOSS_L2 = 0
PTX_LOSS_FD_ISO = 1
PTX_LOSS_FD_ANISO = 2
PTX_LOSS_GRAD_L2 = 3
PTX_LOSS_SPARSE_ISO = 4
PTX_LOSS_SPARSE_ANISO = 5
PTX_ADAM_PROJECT_REAL_NONNEG = 1

_STATUS = {
    0: "PTX_OK",
    1: "PTX_ERR_INVALID_ARGUMENT",
    2: "PTX_ERR_UNSUPPORTED",
    3: "PTX_ERR_OUT_OF_RANGE",
    4: "PTX_ERR_WORKSPACE",
    5: "PTX_ERR_HIP",
}

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t
_u = ctypes.c_uint
_f = ctypes.c_float
_i = ctypes.c_int
_st = ctypes.c_int

# name -> (restype, argtypes); mirrors include/ptx.h
SIGNATURES: dict[str, tuple] = {
    "ptx_last_error": (ctypes.c_char_p, []),
    "ptx_version": (ctypes.c_char_p, []),
    "ptx_abi_version": (_i, []),
    "ptx_timing_events": (_st, [_p, _i]),
    "ptx_timing_count": (_i, []),
    "ptx_timing_label": (ctypes.c_char_p, [_i]),
    "ptx_tune_set": (_st, [_i, _i64, _i]),
    "ptx_tune_get": (_i, [_i, _i64]),
    "ptx_tune_set_groups": (_st, [_i, _i64, _i]),
    "ptx_tune_get_groups": (_i, [_i, _i64]),
    "ptx_stream_copy": (_st, [_p, _p, _sz, _p]),
    "ptx_fft2d": (_st, [_p, _i64, _i64, _i64, _i, _p]),
    "ptx_base_forward": (_st, [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _p, _p]),
    "ptx_base_backward_workspace": (_sz, [_i64, _i64, _i64, _i64, _i64, _u]),
    "ptx_base_backward": (
        _st,
        [_p, _p, _p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_base_step_workspace": (_sz, [_i64, _i64, _i64, _i64, _i64, _u]),
    "ptx_base_step": (
        _st,
        [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _f, _p, _p, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_base_step_indexed": (
        _st,
        [_p, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _p, _i64, _f, _p, _p, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_base_step_noalloc": (
        _st,
        [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _f, _p, _p, _p, _p, _u, _p],
    ),
    "ptx_adam_step": (_st, [_p, _p, _p, _p, _i64, _f, _f, _p, _u, _p]),
    "ptx_base_train_step": (
        _st,
        [_p, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _p, _i64, _f, _p, _p, _p, _p, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_loss_workspace": (_sz, [_i64, _i64, _i64]),
    "ptx_amplitude_loss": (_st, [_i, _p, _p, _i64, _i64, _i64, _f, _f, _p, _p, _p, _sz, _p]),
    "ptx_field_regularizer": (_st, [_i, _p, _i64, _i64, _i64, _f, _f, _p, _p, _p, _sz, _p]),
    "ptx_gather_positions": (_st, [_p, _p, _i64, _i64, _p, _p, _p]),
    "ptx_fpm_forward": (_st, [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _p, _p, _p]),
    "ptx_fpm_backward_workspace": (_sz, [_i64, _i64, _i64, _i64, _i64, _u]),
    "ptx_fpm_backward": (
        _st,
        [_p, _p, _p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_fpm_step_workspace": (_sz, [_i64, _i64, _i64, _i64, _i64, _u]),
    "ptx_fpm_step": (
        _st,
        [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _f, _p, _p, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_nearfield_forward": (_st, [_p, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _p, _p, _p]),
    "ptx_nearfield_backward_workspace": (_sz, [_i64, _i64, _i64, _i64, _i64, _u]),
    "ptx_nearfield_backward": (
        _st,
        [_p, _p, _p, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_nearfield_step_workspace": (_sz, [_i64, _i64, _i64, _i64, _i64, _u]),
    "ptx_nearfield_step": (
        _st,
        [_p, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _p, _f, _p, _p, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_coded_workspace": (_sz, [_i64, _i64, _i64, _i64, _i]),
    "ptx_coded_forward": (_st, [_p, _i64, _i64, _p, _p, _p, _p, _i64, _i64, _p, _p, _p, _p, _p, _sz, _p]),
    "ptx_coded_backward": (
        _st,
        [_p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _p, _p, _p, _sz, _u, _p],
    ),
    "ptx_coded_step": (
        _st,
        [_p, _i64, _i64, _p, _p, _p, _p, _i64, _i64, _p, _f, _p, _p, _p, _p, _p, _sz, _u, _p],
    ),
}



class PtxAdam(ctypes.Structure):
    """``ptx_adam`` of include/ptx.h: one unknown's Adam moments and step-size table."""

    _fields_ = [("exp_avg", _p), ("exp_avg_sq", _p), ("hyper", _p), ("beta1", _f), ("beta2", _f)]


_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


class NativeLibraryError(RuntimeError):
    """The HIP library is missing, failed to load, or a call returned an error."""


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise NativeLibraryError(
                    f"{LIB_PATH.name} is not built; run `python -m ptychox_amd.build` (hipcc, gfx950)"
                )
            try:
                handle = ctypes.CDLL(str(LIB_PATH))
            except OSError as exc:  # pragma: no cover - depends on the runtime
                raise NativeLibraryError(f"cannot load {LIB_PATH}: {exc}") from exc
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def last_error() -> str:
    return lib().ptx_last_error().decode(errors="replace")


def call(name: str, *args) -> None:
    """Call a status-returning entry point; raise with the library's message on error."""
    st = getattr(lib(), name)(*args)
    if st != 0:
        msg = last_error()
        code = _STATUS.get(st, str(st))
        if st in (1, 3):
            raise ValueError(f"{name}: {msg} ({code})")
        raise NativeLibraryError(f"{name}: {msg} ({code})")


def query(name: str, *args) -> int:
    return int(getattr(lib(), name)(*args))


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def stream_of(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


_HIP_EVENT_DISABLE_SYSTEM_FENCE = 0x20000000  # hip_runtime_api.h


def _hip() -> ctypes.CDLL:
    """The HIP runtime already mapped by torch (same soname, one runtime)."""
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipEventCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    h.hipEventDestroy.argtypes = [ctypes.c_void_p]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    return h


class KernelTimer:
    """Per-kernel durations of the next entry-point calls, from hipEvents the
    library records on the call's stream between its kernels
    (``ptx_timing_events``).  The events are created with
    ``hipEventDisableSystemFence``: a timing-only event then adds no system-scope
    cache writeback between kernels, which would otherwise slow the next kernel
    and inflate its measured duration.  Use::

        with KernelTimer(n_events=64) as kt:
            step(...)
        kt.durations()   # [(label, ms), ...] in launch order
    """

    def __init__(self, n_events: int = 64):
        self._h = _hip()
        torch.cuda.synchronize()
        self._ev = []
        for _ in range(n_events):
            e = ctypes.c_void_p()
            if self._h.hipEventCreateWithFlags(ctypes.byref(e), _HIP_EVENT_DISABLE_SYSTEM_FENCE) != 0:
                raise NativeLibraryError("hipEventCreateWithFlags failed")
            self._ev.append(e)
        self._arr = (ctypes.c_void_p * n_events)(*[e.value for e in self._ev])
        self.count, self.labels = 0, []

    def __enter__(self):
        call("ptx_timing_events", ctypes.cast(self._arr, ctypes.c_void_p), len(self._ev))
        return self

    def __exit__(self, *exc):
        self.count = lib().ptx_timing_count()
        self.labels = [lib().ptx_timing_label(k).decode() for k in range(self.count)]
        call("ptx_timing_events", None, 0)
        return False

    def durations(self) -> list[tuple[str, float]]:
        if self.count:
            self._h.hipEventSynchronize(self._ev[self.count - 1])
        out = []
        for k in range(1, self.count):
            ms = ctypes.c_float()
            self._h.hipEventElapsedTime(ctypes.byref(ms), self._ev[k - 1], self._ev[k])
            out.append((self.labels[k], float(ms.value)))
        return out

    def __del__(self):
        for e in getattr(self, "_ev", []):
            self._h.hipEventDestroy(e)

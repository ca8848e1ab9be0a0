"""Drop-in replacements of the reference's native modules
``ptychoX.{base,fpm,coded}._C`` (pybind11 over CUDA + cuFFT:
python/ptychoX/csrc/{base,fpm,coded}_extension.cpp), written as ctypes
bindings of this library's C ABI (include/ptx.h).

The reference loads each module with ``importlib.import_module`` before it
would try to compile the CUDA sources (python/ptychoX/_utils.py:66-76), so a
maintainer makes the reference run on this library by adding one file per
family, e.g. ``python/ptychoX/base/_C.py``::

    from ptychox_amd.ref_c.base import forward, backward, clear_plan_cache

Each module keeps its ``_C`` counterpart's argument order, returned lists and
error behaviour: argument checks raise ``RuntimeError`` with the extension's
``TORCH_CHECK`` messages, a failing library call raises ``RuntimeError`` with
the library's message.  Outputs are fresh tensors from torch's caching
allocator; work runs on the current stream without synchronising.  Where the
reference returns an intermediate that this library recomputes instead of
storing (the conventional patch stack, the FPM sub-spectra), the list slot
carries what the matching backward needs (the sample, the object spectrum);
callers only pass it back to ``backward``, as the reference's operators do.
"""

from __future__ import annotations

import ctypes

import torch

from .. import _lib

_DTYPES = {"complex64": torch.complex64, "float32": torch.float32, "int64": torch.int64}


def check(t: torch.Tensor, name: str, dtype: str) -> None:
    """The extension's TORCH_CHECKs: a CUDA tensor of the given dtype."""
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if t.dtype != _DTYPES[dtype]:
        raise RuntimeError(f"{name} must be {dtype}")


def invoke(name: str, *args) -> None:
    """Call a status-returning entry point; RuntimeError with the library's message on failure."""
    h = _lib.lib()
    st = getattr(h, name)(*args)
    if st != 0:
        raise RuntimeError(f"{name}: {h.ptx_last_error().decode(errors='replace')}")


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def stream(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def scratch(nbytes: int, device: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)

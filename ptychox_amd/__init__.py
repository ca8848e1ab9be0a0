"""MI355X-native (HIP, gfx950) ptychography forward/adjoint operators.

Drop-in for the reference package ``ptychoX`` (python/ptychoX/__init__.py):
the same operator names, signatures, aliases and cache-clearing calls, plus a
near-field operator.  ``use_cuda_kernels=True`` on complex64 GPU tensors runs
the hand-written HIP kernels of ``libptychox_hip.so``.
"""

from ._utils import matlab_positions_to_torch, recompute_mode, scatter_mode, set_recompute, set_scatter_mode
from .base import clear_cuda_plan_cache as clear_base_cuda_plan_cache
from .base import ptychoX_base, ptychography_base
from .coded import clear_cuda_plan_cache as clear_coded_cuda_plan_cache
from .coded import coded_ptychography, ptychoX_coded
from .fpm import clear_cuda_plan_cache as clear_fpm_cuda_plan_cache
from .fpm import fourier_ptychography, ptychoX_fourier
from .nearfield import angular_spectrum_transfer, nearfield_ptychography, ptychoX_nearfield
from .nearfield import clear_cuda_plan_cache as clear_nearfield_cuda_plan_cache
from .nearfield_matlab import AdPtychoXNearf, ptychoX_nearfield_matlab


def clear_cuda_plan_cache() -> None:
    """Kept for drop-in compatibility (python/ptychoX/__init__.py:10-13)."""
    clear_base_cuda_plan_cache()
    clear_fpm_cuda_plan_cache()
    clear_coded_cuda_plan_cache()
    clear_nearfield_cuda_plan_cache()


def install_as_ptychoX() -> None:
    """Make this package importable under the reference's name, so code
    written against it runs unchanged: ``import ptychoX``, ``from ptychoX.ops
    import *`` (python/ops.py:3) and ``ptychoX.{base,fpm,coded}.ops`` resolve to
    this package and its modules.  Refuses (ImportError) when a different
    ``ptychoX`` is already imported."""
    import sys

    from . import base, coded, fpm, nearfield, ops

    pkg = sys.modules[__name__]
    cur = sys.modules.get("ptychoX")
    if cur is not None and cur is not pkg:
        raise ImportError("another package named ptychoX is already imported")
    sys.modules["ptychoX"] = pkg
    sys.modules["ptychoX.ops"] = ops
    for name, mod in (("base", base), ("fpm", fpm), ("coded", coded), ("nearfield", nearfield)):
        sys.modules[f"ptychoX.{name}"] = mod
        sys.modules[f"ptychoX.{name}.ops"] = mod
        if not hasattr(mod, "ops"):  # `import ptychoX.base.ops as m` reads the attribute
            mod.ops = mod


__all__ = [
    "clear_cuda_plan_cache",
    "clear_base_cuda_plan_cache",
    "clear_fpm_cuda_plan_cache",
    "clear_coded_cuda_plan_cache",
    "coded_ptychography",
    "ptychoX_coded",
    "clear_nearfield_cuda_plan_cache",
    "angular_spectrum_transfer",
    "nearfield_ptychography",
    "ptychoX_nearfield",
    "AdPtychoXNearf",
    "ptychoX_nearfield_matlab",
    "fourier_ptychography",
    "ptychoX_fourier",
    "matlab_positions_to_torch",
    "ptychoX_base",
    "ptychography_base",
    "scatter_mode",
    "set_scatter_mode",
    "recompute_mode",
    "set_recompute",
    "install_as_ptychoX",
]

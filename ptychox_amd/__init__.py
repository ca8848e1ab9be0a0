"""MI355X-native (HIP, gfx950) ptychography forward/adjoint operators.

Drop-in for the reference package ``ptychoX`` (python/ptychoX/__init__.py):
the same operator names, signatures, aliases and cache-clearing calls, plus a
near-field operator.  ``use_cuda_kernels=True`` on complex64 GPU tensors runs
the hand-written HIP kernels of ``libptychox_hip.so``.
"""

from ._utils import matlab_positions_to_torch, set_deterministic
from .base import clear_cuda_plan_cache as clear_base_cuda_plan_cache
from .base import ptychoX_base, ptychography_base


def clear_cuda_plan_cache() -> None:
    """Kept for drop-in compatibility (python/ptychoX/__init__.py:10-13)."""
    clear_base_cuda_plan_cache()


__all__ = [
    "clear_cuda_plan_cache",
    "clear_base_cuda_plan_cache",
    "matlab_positions_to_torch",
    "ptychoX_base",
    "ptychography_base",
    "set_deterministic",
]

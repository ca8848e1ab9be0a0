"""Convenience imports of the operators, under the names the reference's
``ptychoX.ops`` exports (python/ptychoX/ops.py:3-14), plus the near-field
operator.  ``from ptychox_amd.ops import *`` replaces the reference shim
``from ptychoX.ops import *`` (python/ops.py:3); after
``ptychox_amd.install_as_ptychoX()`` that shim line itself works unchanged."""

from . import (  # noqa: F401
    clear_base_cuda_plan_cache,
    clear_coded_cuda_plan_cache,
    clear_cuda_plan_cache,
    clear_fpm_cuda_plan_cache,
    clear_nearfield_cuda_plan_cache,
    coded_ptychography,
    fourier_ptychography,
    matlab_positions_to_torch,
    nearfield_ptychography,
    ptychoX_base,
    ptychoX_coded,
    ptychoX_fourier,
    ptychoX_nearfield,
    ptychography_base,
)

__all__ = [
    "clear_base_cuda_plan_cache",
    "clear_cuda_plan_cache",
    "clear_fpm_cuda_plan_cache",
    "coded_ptychography",
    "fourier_ptychography",
    "matlab_positions_to_torch",
    "ptychoX_base",
    "ptychoX_coded",
    "ptychoX_fourier",
    "ptychography_base",
    "clear_coded_cuda_plan_cache",
    "clear_nearfield_cuda_plan_cache",
    "nearfield_ptychography",
    "ptychoX_nearfield",
]

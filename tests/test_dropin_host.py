"""The drop-in surface without a GPU: ``ptychox_amd.ops`` exports the names of
the reference's ``ptychoX.ops`` (python/ptychoX/ops.py:3-14) with the
reference's signatures, ``install_as_ptychoX`` makes the reference's own import
lines work, the ``_C`` replacement modules keep the extension's argument
checks (python/ptychoX/csrc/*_extension.cpp), and the recompute policy."""

from __future__ import annotations

import inspect
import subprocess
import sys
from pathlib import Path

import pytest
import torch

ROOT = Path(__file__).resolve().parents[1]

# python/ptychoX/ops.py:3-14
REFERENCE_OPS = [
    "clear_base_cuda_plan_cache", "clear_cuda_plan_cache", "clear_fpm_cuda_plan_cache", "coded_ptychography",
    "fourier_ptychography", "matlab_positions_to_torch", "ptychoX_base", "ptychoX_coded", "ptychoX_fourier",
    "ptychography_base",
]

# parameter lists of the reference's operators (base/ops.py:14-19, fpm/ops.py:14-19,
# coded/ops.py:22-30, _utils.py:99)
SIGNATURES = {
    "ptychoX_base": ["sample", "probe", "positions_xy", "use_cuda_kernels"],
    "ptychoX_fourier": ["object_field", "pupil", "led_indices_xy", "use_cuda_kernels"],
    "ptychoX_coded": ["object_field", "coded_surface", "transfer_1", "transfer_2", "shifts_xy", "downsample",
                      "use_cuda_kernels"],
    "matlab_positions_to_torch": ["positions_xy_1based"],
    "clear_cuda_plan_cache": [],
    "clear_base_cuda_plan_cache": [],
    "clear_fpm_cuda_plan_cache": [],
}


def test_ops_exports_the_reference_names():
    from ptychox_amd import ops

    for name in REFERENCE_OPS:
        assert hasattr(ops, name), name
        assert name in ops.__all__
    assert ops.ptychography_base is ops.ptychoX_base
    assert ops.fourier_ptychography is ops.ptychoX_fourier
    assert ops.coded_ptychography is ops.ptychoX_coded
    for name, params in SIGNATURES.items():
        sig = inspect.signature(getattr(ops, name))
        assert list(sig.parameters) == params, name
        if "use_cuda_kernels" in sig.parameters:
            assert sig.parameters["use_cuda_kernels"].default is False
    # the star import gives every reference name
    ns: dict = {}
    exec("from ptychox_amd.ops import *", ns)
    assert set(REFERENCE_OPS) <= set(ns)


def test_reference_import_lines_after_install():
    # in a fresh interpreter: the reference shim line (python/ops.py:3) and the
    # package-internal module paths resolve to this package
    code = (
        "import ptychox_amd\n"
        "ptychox_amd.install_as_ptychoX()\n"
        "from ptychoX.ops import *\n"
        "import ptychoX\n"
        "from ptychoX.base.ops import ptychoX_base as b\n"
        "from ptychoX.fpm.ops import fourier_ptychography as f\n"
        "import ptychoX.coded.ops as c\n"
        "assert b is ptychoX_base and f is fourier_ptychography and c.ptychoX_coded is ptychoX_coded\n"
        "assert ptychoX.matlab_positions_to_torch is matlab_positions_to_torch\n"
        "clear_cuda_plan_cache()\n"
        "print('ok')\n"
    )
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().endswith("ok")


def test_install_refuses_a_foreign_ptychoX():
    code = (
        "import sys, types\n"
        "sys.modules['ptychoX'] = types.ModuleType('ptychoX')\n"
        "import ptychox_amd\n"
        "try:\n"
        "    ptychox_amd.install_as_ptychoX()\n"
        "except ImportError:\n"
        "    print('refused')\n"
    )
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.stdout.strip() == "refused", out.stderr


@pytest.mark.parametrize("family", ["base", "fpm", "coded"])
def test_c_modules_have_the_extension_surface(family):
    import importlib

    mod = importlib.import_module(f"ptychox_amd.ref_c.{family}")
    for fn in ("forward", "backward", "clear_plan_cache"):
        assert callable(getattr(mod, fn))
    mod.clear_plan_cache()
    params = {
        "base": (["sample", "probe", "positions_xy"],
                 ["grad_output", "patch_stack", "probe", "far_field_shifted", "positions_xy", "sample_h",
                  "sample_w"]),
        "fpm": (["object", "pupil", "led_indices_xy"],
                ["grad_output", "sub_spectrum", "pupil", "image_field", "led_indices_xy", "object_h", "object_w"]),
        "coded": (["object", "coded_surface", "transfer_1", "transfer_2", "shifts_xy", "downsample"],
                  ["grad_output", "object_spectrum", "x_forward", "sensor_field", "coded_surface", "transfer_1",
                   "transfer_2", "shifts_xy", "downsample"]),
    }[family]
    assert list(inspect.signature(mod.forward).parameters) == params[0]
    assert list(inspect.signature(mod.backward).parameters) == params[1]


def test_c_modules_check_arguments_like_the_extension():
    from ptychox_amd.ref_c import base, coded, fpm

    c = torch.zeros(8, 8, dtype=torch.complex64)
    pos = torch.zeros(1, 2, dtype=torch.int64)
    with pytest.raises(RuntimeError, match="^sample must be a CUDA tensor$"):
        base.forward(c, c, pos)
    with pytest.raises(RuntimeError, match="^object must be a CUDA tensor$"):
        fpm.forward(c, c, pos)
    with pytest.raises(RuntimeError, match="^object must be a CUDA tensor$"):
        coded.forward(c, c, c, c, torch.zeros(1, 2), 1)
    with pytest.raises(RuntimeError, match="^grad_output must be a CUDA tensor$"):
        base.backward(torch.zeros(1, 8, 8), c, c, c, pos, 8, 8)


def test_recompute_policy():
    from ptychox_amd import _utils as u

    cpu = torch.device("cpu")
    try:
        u.set_recompute("never")
        assert not u.should_recompute(1 << 40, cpu)
        u.set_recompute("always")
        assert u.should_recompute(0, cpu)
        u.set_recompute(1000)
        assert u.should_recompute(1001, cpu) and not u.should_recompute(1000, cpu)
        u.set_recompute(None)
        assert u.recompute_mode() == "auto"
        dev = torch.device("cuda", 0)
        assert u.should_recompute(51, dev, free_bytes=100) and not u.should_recompute(50, dev, free_bytes=100)
        for bad in ("sometimes", -1, True, 1.5):
            with pytest.raises(ValueError):
                u.set_recompute(bad)
    finally:
        u.set_recompute(None)

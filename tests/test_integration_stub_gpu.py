"""The reference-side ctypes binding shown in the integration guide (a drop-in
for the pybind module ptychoX.base._C, python/ptychoX/base/ops.py:110-124), run
as written against the built library: its forward must return what the
package's own ptychoX_base returns."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from tests.helpers import disc_probe, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu


def _stub():
    from ptychox_amd import _lib as pkg_lib

    pkg_lib.lib()  # loads the library (and fails loudly when it is missing)
    lib = ctypes.CDLL(str(pkg_lib.LIB_PATH))
    lib.ptx_base_forward.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ptx_last_error.restype = ctypes.c_char_p

    def forward(sample, probe, positions_xy):  # base_extension.cpp:21-32
        B, (h, w) = positions_xy.shape[0], probe.shape
        far = torch.empty((B, h, w), dtype=torch.complex64, device=sample.device)
        amp = torch.empty((B, h, w), dtype=torch.float32, device=sample.device)
        st = lib.ptx_base_forward(sample.data_ptr(), *sample.shape, probe.data_ptr(), h, w,
                                  positions_xy.data_ptr(), B, far.data_ptr(), amp.data_ptr(),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        if st != 0:
            raise RuntimeError(lib.ptx_last_error().decode())
        return [sample, far, amp]

    return forward, lib


def test_reference_side_ctypes_stub(gpu):
    from ptychox_amd import ptychoX_base

    forward, lib = _stub()
    rng = np.random.default_rng(8)
    s = to_dev(smooth_object(rng, 160), gpu)
    p = to_dev(disc_probe(64, 16), gpu)
    pos = to_dev(rng.integers(0, 97, (5, 2)).astype(np.int64), gpu)
    _, far, amp = forward(s, p, pos)
    torch.cuda.synchronize()
    ref = ptychoX_base(s, p, pos, use_cuda_kernels=True)
    assert torch.equal(amp, ref)
    assert nrel(far.abs(), ref) < 1e-6
    # positions are not checked by the library: a window outside the sample
    # yields a zero frame and leaves the other views unchanged
    bad = pos.clone()
    bad[0, 0] = 1000
    _, _, amp_bad = forward(s, p, bad)
    torch.cuda.synchronize()
    assert not bool(amp_bad[0].any())
    assert torch.equal(amp_bad[1:], ref[1:])
    # a probe larger than the sample is an error status with a message
    big = to_dev(disc_probe(256, 64), gpu)
    with pytest.raises(RuntimeError, match="larger than sample"):
        forward(s, big, pos)

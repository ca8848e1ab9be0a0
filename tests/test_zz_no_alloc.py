"""ptx_base_step_noalloc (scratch in a static device array) against
ptx_base_step (caller workspace), through the C ABI: two calls in a row, the
second on a smaller problem (so the array holds the first call's data when the
second starts), in both gradient modes.  Same kernels, same scratch layout:
results are bitwise equal in the sorted-segment mode; with atomics the object
gradient's summation order varies, so it is held to 1e-6."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.helpers import crandn, disc_probe, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

# (object side, probe side, batch): the headline shape first, then a smaller one
CALLS = [(2048, 512, 64), (512, 128, 20)]


def _problem(gpu, H, h, B, seed):
    rng = np.random.default_rng(seed)
    s = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    p = disc_probe(h, h / 4)
    pos = rng.integers(0, H - h + 1, (B, 2)).astype(np.int64)
    y = rng.uniform(0.0, 3.0, (B, h, h)).astype(np.float32)
    return [to_dev(a, gpu) for a in (s, p, pos, y)]


def _run(gpu, name, s, p, pos, y, flags):
    from ptychox_amd import _lib

    H, W = s.shape
    h, w = p.shape
    B = pos.shape[0]
    loss = torch.empty(B, dtype=torch.float32, device=gpu)
    amp = torch.empty((B, h, w), dtype=torch.float32, device=gpu)
    gs, gp = torch.empty_like(s), torch.empty_like(p)
    scale = 1.0 / (B * h * w)
    head = (_lib.ptr(s), H, W, _lib.ptr(p), h, w, _lib.ptr(pos), B, _lib.ptr(y), scale, _lib.ptr(loss),
            _lib.ptr(amp), _lib.ptr(gs), _lib.ptr(gp))
    if name == "ptx_base_step":
        ws = _lib.workspace(_lib.query("ptx_base_step_workspace", H, W, h, w, B, flags), gpu)
        _lib.call(name, *head, _lib.ptr(ws), ws.numel(), flags, _lib.stream_of(gpu))
    else:
        _lib.call(name, *head, flags, _lib.stream_of(gpu))
    torch.cuda.synchronize()
    return loss, amp, gs, gp


@pytest.mark.parametrize("mode", ["segments", "atomic"])
def test_noalloc_step_matches_workspace_step(gpu, mode):
    from ptychox_amd._utils import scatter_flags

    flags = scatter_flags(mode)
    for k, (H, h, B) in enumerate(CALLS):
        print(f"noalloc vs workspace step: mode {mode}, object {H}, probe {h}, batch {B}", flush=True)
        args = _problem(gpu, H, h, B, seed=100 + k)
        ref = _run(gpu, "ptx_base_step", *args, flags)
        got = _run(gpu, "ptx_base_step_noalloc", *args, flags)
        loss, amp, gs, gp = got
        assert torch.equal(loss, ref[0])
        assert torch.equal(amp, ref[1])
        assert torch.equal(gp, ref[3])
        if mode == "segments":
            assert torch.equal(gs, ref[2])
        else:
            assert nrel(gs, ref[2]) < 1e-6


def test_noalloc_step_rejects_oversize(gpu):
    from ptychox_amd import _lib

    # 2048^2 / 512^2 with 128 views needs ~290 MiB of scratch: refused before any launch
    args = _problem(gpu, 2048, 512, 128, seed=7)
    with pytest.raises(_lib.NativeLibraryError, match="static maximum"):
        _run(gpu, "ptx_base_step_noalloc", *args, 0)

"""A scan window that leaves the sample contributes nothing to the gradients.

Positions are device data and the C ABI does not check them: such a view gives
a zero frame and, in the adjoint and step entry points, no gradient
(include/ptx.h; the reference's CUDA kernel would read out of bounds,
python/ptychoX/csrc/base_kernel.cu:108-137).  In the default sorted-segment
mode the scatter writes no window contribution for such a view, so its rows of
the contribution stack keep whatever the column pass left there -- the tile
lists must not list it.  Each test runs a batch with one out-of-range view
(past the right edge, or at a negative corner) and the same batch without it:
with a caller-supplied non-zero far field for that view (the backward) or real
measurements (the steps), the gradients must be bitwise equal in the
sorted-segment mode and equal to 1e-6 with atomics.  The probe gradient is
compared bitwise with one probe-partial slice (work-group target 1), so both
batches sum their views in the same order.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.helpers import crandn, disc_probe, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

H, h, B, K = 256, 64, 12, 5
BAD = {"right": (H - h + 9, 20), "negative": (-3, 40), "below": (30, H - h + 1)}


@pytest.fixture
def one_probe_slice():
    from ptychox_amd import _lib

    lib = _lib.lib()
    old = lib.ptx_tune_get_groups(1, h)
    _lib.call("ptx_tune_set_groups", 1, h, 1)
    yield
    _lib.call("ptx_tune_set_groups", 1, h, old)


def _case(seed, where):
    rng = np.random.default_rng(seed)
    s = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    p = disc_probe(h, h / 4)
    pos = rng.integers(0, H - h + 1, (B, 2)).astype(np.int64)
    pos[K] = BAD[where]
    keep = np.array([b for b in range(B) if b != K])
    return rng, s, p, pos, keep


def _compare(mode, with_bad, without):
    gs1, gp1 = with_bad
    gs0, gp0 = without
    if mode == "segments":
        assert torch.equal(gs1, gs0)
        assert torch.equal(gp1, gp0)
    else:
        assert nrel(gs1, gs0) < 1e-6
        assert torch.equal(gp1, gp0)


@pytest.mark.parametrize("mode", ["segments", "atomic"])
@pytest.mark.parametrize("where", list(BAD))
def test_base_backward_skips_view(gpu, one_probe_slice, mode, where):
    from ptychox_amd.base import native_backward
    from ptychox_amd._utils import scatter_flags

    rng, s, p, pos, keep = _case(11, where)
    g = rng.uniform(0.1, 1.0, (B, h, h)).astype(np.float32)
    far = crandn(rng, B, h, h)  # non-zero for the out-of-range view too
    sd, pd = to_dev(s, gpu), to_dev(p, gpu)
    flags = scatter_flags(mode)
    with_bad = native_backward(to_dev(g, gpu), to_dev(far, gpu), sd, pd, to_dev(pos, gpu), flags)
    without = native_backward(to_dev(g[keep], gpu), to_dev(far[keep], gpu), sd, pd, to_dev(pos[keep], gpu), flags)
    torch.cuda.synchronize()
    _compare(mode, with_bad, without)


def _measured(rng, n):
    return rng.uniform(0.0, 3.0, (n, h, h)).astype(np.float32)


@pytest.mark.parametrize("mode", ["segments", "atomic"])
@pytest.mark.parametrize("where", list(BAD))
def test_base_step_skips_view(gpu, one_probe_slice, mode, where):
    from ptychox_amd.base import BaseStep

    rng, s, p, pos, keep = _case(12, where)
    y = _measured(rng, B)
    sd, pd = to_dev(s, gpu), to_dev(p, gpu)
    scale = 1.0 / (B * h * h)
    res = []
    for pp, yy in ((pos, y), (pos[keep], y[keep])):
        step = BaseStep(s.shape, p.shape, len(pp), gpu, scatter=mode, keep_amplitude=True)
        gs = torch.empty_like(sd)
        gp = torch.empty_like(pd)
        loss = step(sd, pd, to_dev(pp, gpu), to_dev(yy, gpu), gs, gp, loss_scale=scale).clone()
        torch.cuda.synchronize()
        res.append((gs, gp, loss, step.amplitude.clone()))
    (gs1, gp1, l1, a1), (gs0, gp0, l0, a0) = res
    _compare(mode, (gs1, gp1), (gs0, gp0))
    assert torch.equal(l1[keep], l0)
    assert torch.equal(a1[keep], a0)
    assert not bool(a1[K].any())
    # the out-of-range view's loss: its frame is zero, so mean(Y^2)
    assert abs(float(l1[K]) - float(np.mean(y[K].astype(np.float64) ** 2))) < 1e-5 * float(l1[K])


@pytest.mark.parametrize("mode", ["segments", "atomic"])
def test_base_step_indexed_skips_view(gpu, one_probe_slice, mode):
    from ptychox_amd import _lib
    from ptychox_amd._utils import scatter_flags

    rng = np.random.default_rng(13)
    s = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    p = disc_probe(h, h / 4)
    count = 30
    pos_all = rng.integers(0, H - h + 1, (count, 2)).astype(np.int64)
    pos_all[17] = BAD["right"]
    y_all = _measured(rng, count)
    order = rng.permutation(count)[:B]
    order[K] = 17
    keep = np.array([b for b in range(B) if b != K])
    sd, pd = to_dev(s, gpu), to_dev(p, gpu)
    pa, ya = to_dev(pos_all, gpu), to_dev(y_all, gpu)
    flags = scatter_flags(mode)
    scale = 1.0 / (B * h * h)
    res = []
    for od in (order, order[keep]):
        n = len(od)
        ws = _lib.workspace(_lib.query("ptx_base_step_workspace", H, H, h, h, n, flags), gpu)
        gs, gp = torch.empty_like(sd), torch.empty_like(pd)
        loss = torch.empty(n, dtype=torch.float32, device=gpu)
        _lib.call("ptx_base_step_indexed", _lib.ptr(sd), H, H, _lib.ptr(pd), h, h, _lib.ptr(pa), _lib.ptr(ya),
                  count, _lib.ptr(to_dev(od.astype(np.int64), gpu)), n, scale, _lib.ptr(loss), None,
                  _lib.ptr(gs), _lib.ptr(gp), _lib.ptr(ws), ws.numel(), flags, _lib.stream_of(gpu))
        torch.cuda.synchronize()
        res.append((gs, gp, loss))
    (gs1, gp1, l1), (gs0, gp0, l0) = res
    _compare(mode, (gs1, gp1), (gs0, gp0))
    assert torch.equal(l1[keep], l0)


@pytest.mark.parametrize("mode", ["segments", "atomic"])
@pytest.mark.parametrize("where", ["right", "negative"])
def test_nearfield_skips_view(gpu, one_probe_slice, mode, where):
    from ptychox_amd.nearfield import NearFieldStep, angular_spectrum_transfer, native_backward
    from ptychox_amd._utils import scatter_flags

    rng, s, p, pos, keep = _case(14, where)
    sd, pd = to_dev(s, gpu), to_dev(p, gpu)
    tr = angular_spectrum_transfer(h, h, (h / 4) ** 2 / 20.0, device=gpu)
    flags = scatter_flags(mode)
    # backward with a non-zero field for the out-of-range view
    g = rng.uniform(0.1, 1.0, (B, h, h)).astype(np.float32)
    field = crandn(rng, B, h, h)
    with_bad = native_backward(to_dev(g, gpu), to_dev(field, gpu), sd, pd, tr, to_dev(pos, gpu), flags)
    without = native_backward(to_dev(g[keep], gpu), to_dev(field[keep], gpu), sd, pd, tr, to_dev(pos[keep], gpu),
                              flags)
    torch.cuda.synchronize()
    _compare(mode, with_bad, without)
    # fused step
    y = _measured(rng, B)
    scale = 1.0 / (B * h * h)
    res = []
    for pp, yy in ((pos, y), (pos[keep], y[keep])):
        step = NearFieldStep(s.shape, p.shape, len(pp), gpu, scatter=mode)
        gs, gp = torch.empty_like(sd), torch.empty_like(pd)
        loss = step(sd, pd, tr, to_dev(pp, gpu), to_dev(yy, gpu), gs, gp, loss_scale=scale).clone()
        torch.cuda.synchronize()
        res.append((gs, gp, loss))
    (gs1, gp1, l1), (gs0, gp0, l0) = res
    _compare(mode, (gs1, gp1), (gs0, gp0))
    assert torch.equal(l1[keep], l0)


@pytest.mark.parametrize("side", [64, 32])  # three-pass path (positions gathered in the gather rows), whole frame
@pytest.mark.parametrize("mode", ["segments", "atomic"])
def test_base_step_indexed_order_outside_table(gpu, mode, side):
    # an order entry outside [0, count) reads entry 0 of both tables (include/ptx.h):
    # bitwise the indexed step whose order holds 0 there
    from ptychox_amd import _lib
    from ptychox_amd._utils import scatter_flags

    rng = np.random.default_rng(21 + side)
    Hs = 4 * side
    s = (smooth_object(rng, Hs) + 0.05 * crandn(rng, Hs, Hs)).astype(np.complex64)
    p = disc_probe(side, side / 4)
    count, n = 30, 12
    pos_all = rng.integers(0, Hs - side + 1, (count, 2)).astype(np.int64)
    y_all = rng.uniform(0.0, 2.0, (count, side, side)).astype(np.float32)
    order = rng.permutation(count)[:n].astype(np.int64)
    bad = order.copy()
    bad[3], bad[7] = -1, count + 5
    good = order.copy()
    good[3], good[7] = 0, 0
    sd, pd, pa, ya = to_dev(s, gpu), to_dev(p, gpu), to_dev(pos_all, gpu), to_dev(y_all, gpu)
    flags = scatter_flags(mode)
    ws = _lib.workspace(_lib.query("ptx_base_step_workspace", Hs, Hs, side, side, n, flags), gpu)
    res = []
    for od in (bad, good):
        gs, gp = torch.empty_like(sd), torch.empty_like(pd)
        loss = torch.empty(n, dtype=torch.float32, device=gpu)
        _lib.call("ptx_base_step_indexed", _lib.ptr(sd), Hs, Hs, _lib.ptr(pd), side, side, _lib.ptr(pa),
                  _lib.ptr(ya), count, _lib.ptr(to_dev(od, gpu)), n, 1.0 / (n * side * side), _lib.ptr(loss), None,
                  _lib.ptr(gs), _lib.ptr(gp), _lib.ptr(ws), ws.numel(), flags, _lib.stream_of(gpu))
        torch.cuda.synchronize()
        res.append((gs, gp, loss))
    (gs1, gp1, l1), (gs0, gp0, l0) = res
    assert torch.equal(l1, l0)
    if mode == "segments":
        assert torch.equal(gs1, gs0) and torch.equal(gp1, gp0)
    else:
        assert nrel(gs1, gs0) < 1e-6 and nrel(gp1, gp0) < 1e-6

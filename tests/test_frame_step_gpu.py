"""The whole-frame fused conventional step (square probes of side 32, 64 and
128, PTX_FLAG_WHOLE_FRAME; the default for 32: one work-group per chunk of
views with the frame in LDS) against the FP64
oracle (python/train_ptychography.py:549-558 restated: ptychoX_base ->
F.mse_loss -> backward) and against the three-pass path on the same inputs
(PTX_FLAG_THREE_PASS), in both gradient modes, with amplitudes, chunks of
several views per work-group, an out-of-range view, accumulation and the
indexed (permuted mini-batch) entry point.  Bars: norm-wise relative error
<= 1e-5 against the oracle; <= 2e-6 between the two paths."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import crandn, disc_probe, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture
def groups_target():
    """Set the work-group target of the view-walking row passes (which fixes
    the chunk of views per work-group) for one probe side; restored after."""
    from ptychox_amd import _lib

    saved = []

    def set_(n, groups):
        saved.append((n, _lib.lib().ptx_tune_get_groups(1, n)))
        _lib.call("ptx_tune_set_groups", 1, n, groups)

    yield set_
    for n, g in reversed(saved):
        _lib.call("ptx_tune_set_groups", 1, n, g)


def _case(seed, H, h, B):
    rng = np.random.default_rng(seed)
    s = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    p = (disc_probe(h, h / 4) + 0.02 * crandn(rng, h, h)).astype(np.complex64)
    pos = rng.integers(0, H - h + 1, (B, 2)).astype(np.int64)
    y = po.base_forward(s * 1.1, p, pos)["amplitude"].astype(np.float32)
    return s, p, pos, y


def _run(gpu, s, p, pos, y, scatter, three_pass, accumulate=False, g0=None):
    from ptychox_amd.base import BaseStep

    B = pos.shape[0]
    step = BaseStep(s.shape, p.shape, B, gpu, scatter=scatter, keep_amplitude=True,
                    path="three_pass" if three_pass else "whole_frame")
    gs = torch.zeros(s.shape, dtype=torch.complex64, device=gpu) if g0 is None else g0.clone()
    gp = torch.zeros(p.shape, dtype=torch.complex64, device=gpu)
    loss = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos, gpu), to_dev(y, gpu), gs, gp,
                accumulate=accumulate).clone()
    torch.cuda.synchronize()
    return loss, step.amplitude.clone(), gs, gp


@pytest.mark.parametrize("h,H,B", [(32, 80, 7), (64, 160, 12), (128, 320, 9), (128, 512, 40), (64, 96, 1)])
@pytest.mark.parametrize("scatter", ["segments", "atomic"])
def test_frame_step_vs_oracle_and_three_pass(gpu, h, H, B, scatter):
    s, p, pos, y = _case(h + B, H, h, B)
    _, per, gs_ref, gp_ref = po.base_train_step(s, p, pos, y)
    amp_ref = po.base_forward(s, p, pos)["amplitude"]
    frame = _run(gpu, s, p, pos, y, scatter, three_pass=False)
    three = _run(gpu, s, p, pos, y, scatter, three_pass=True)
    e_or = (nrel(frame[0], per), nrel(frame[1], amp_ref), nrel(frame[2], gs_ref), nrel(frame[3], gp_ref))
    e_3p = tuple(nrel(a, b) for a, b in zip(frame, three))
    print(f"frame step {H}/{h}/B{B} {scatter}: vs FP64 loss_b {e_or[0]:.2e} amp {e_or[1]:.2e} grad_sample "
          f"{e_or[2]:.2e} grad_probe {e_or[3]:.2e}; vs three-pass {max(e_3p):.2e}", flush=True)
    assert max(e_or) < TOL
    assert max(e_3p) < 2e-6


@pytest.mark.parametrize("h", [32, 128])
def test_frame_step_chunks_of_views(gpu, groups_target, h):
    # work-group target 1: one work-group walks all views (chunk = batch);
    # the probe partial of the chunk accumulates in its slice
    s, p, pos, y = _case(3 * h, 3 * h, h, 11)
    _, per, gs_ref, gp_ref = po.base_train_step(s, p, pos, y)
    groups_target(h, 1)
    loss, _, gs, gp = _run(gpu, s, p, pos, y, "segments", three_pass=False)
    e = (nrel(loss, per), nrel(gs, gs_ref), nrel(gp, gp_ref))
    print(f"frame step chunk = batch, h {h}: {max(e):.2e}", flush=True)
    assert max(e) < TOL


@pytest.mark.parametrize("scatter", ["segments", "atomic"])
def test_frame_step_out_of_range_and_accumulate(gpu, groups_target, scatter):
    h, H, B = 64, 192, 10
    s, p, pos, y = _case(5, H, h, B)
    bad = pos.copy()
    bad[3] = [H - h + 5, 7]  # leaves the sample: zero frame, no gradient
    keep = np.array([b for b in range(B) if b != 3])
    groups_target(h, 8)  # chunks of several views, one of them starting with the bad view
    g0 = to_dev(crandn(np.random.default_rng(1), H, H), gpu)
    loss, amp, gs, gp = _run(gpu, s, p, bad, y, scatter, three_pass=False, accumulate=True, g0=g0)
    # the batch without the bad view, same loss scale
    from ptychox_amd.base import BaseStep

    step = BaseStep(s.shape, p.shape, B - 1, gpu, scatter=scatter, keep_amplitude=True, path="whole_frame")
    gs0, gp0 = g0.clone(), torch.zeros(p.shape, dtype=torch.complex64, device=gpu)
    l0 = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos[keep], gpu), to_dev(y[keep], gpu), gs0, gp0,
              loss_scale=1.0 / (B * h * h), accumulate=True).clone()
    torch.cuda.synchronize()
    assert not bool(amp[3].any())
    assert nrel(loss[keep], l0) < 1e-6  # per-position losses are frame means
    assert nrel(gs, gs0) < 1e-6
    assert nrel(gp, gp0) < 1e-6


def test_frame_step_indexed(gpu):
    from ptychox_amd import _lib

    h, H, count, B = 128, 384, 30, 13
    s, p, pos_all, y_all = _case(9, H, h, count)
    rng = np.random.default_rng(2)
    order = rng.permutation(count)[:B].astype(np.int64)
    _, per, gs_ref, gp_ref = po.base_train_step(s, p, pos_all[order], y_all[order])
    sd, pd = to_dev(s, gpu), to_dev(p, gpu)
    flags = _lib.PTX_FLAG_WHOLE_FRAME
    ws = _lib.workspace(_lib.query("ptx_base_step_workspace", H, H, h, h, B, flags), gpu)
    gs, gp = torch.empty_like(sd), torch.empty_like(pd)
    loss = torch.empty(B, dtype=torch.float32, device=gpu)
    _lib.call("ptx_base_step_indexed", _lib.ptr(sd), H, H, _lib.ptr(pd), h, h, _lib.ptr(to_dev(pos_all, gpu)),
              _lib.ptr(to_dev(y_all, gpu)), count, _lib.ptr(to_dev(order, gpu)), B, 1.0 / (B * h * h),
              _lib.ptr(loss), None, _lib.ptr(gs), _lib.ptr(gp), _lib.ptr(ws), ws.numel(), flags, _lib.stream_of(gpu))
    torch.cuda.synchronize()
    e = (nrel(loss, per), nrel(gs, gs_ref), nrel(gp, gp_ref))
    print(f"frame step indexed: {max(e):.2e}", flush=True)
    assert max(e) < TOL

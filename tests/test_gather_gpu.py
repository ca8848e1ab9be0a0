"""The deterministic object/spectrum gradient gather in both of its modes.

Windows that overlap by fewer than eight views per pixel on average are
gathered in one pass per 4 x 128 tile; denser overlaps split the view list of
every 16 x 128 tile into segments of eight views, gathered in parallel and
combined in segment order.  Each mode is checked against the FP64 oracle (norm-wise
relative error <= 1e-5, FP32), against the atomic scatter, for run-to-run
bitwise equality, and with gradient accumulation (PTX_FLAG_ACCUMULATE).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import crandn, disc_probe, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _segmented(B, h, H):
    return B * h * h >= 8 * H * H


def _base_case(rng, H, h, B, spot):
    s = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    p = disc_probe(h, h / 4)
    if spot:  # every view on (nearly) the same window: one tile list of B views
        pos = np.tile(rng.integers(0, H - h, 2), (B, 1))
        pos[::3] += 1
    else:
        pos = np.stack([rng.integers(0, H - h + 1, B), rng.integers(0, H - h + 1, B)], axis=1)
    return s, p, pos.astype(np.int64)


# (H, h, B, spot): single pass, segmented, segmented with >200-view tiles, segmented on a
# width of partial tiles, single pass on a height of partial tiles, single pass with 70-view
# lists (more than one 64-view list chunk), a batch of one view
CASES = [(256, 64, 9, False), (128, 64, 40, False), (160, 32, 203, True), (200, 64, 80, False),
         (203, 64, 12, False), (512, 64, 70, True), (96, 32, 1, False)]


@pytest.mark.parametrize("H,h,B,spot", CASES)
def test_base_step_gather_modes_vs_oracle(gpu, H, h, B, spot):
    from ptychox_amd.base import BaseStep

    rng = np.random.default_rng(H * 3 + B)
    s, p, pos = _base_case(rng, H, h, B, spot)
    y = np.abs(po.base_forward(s * 1.05, p, pos)["amplitude"]).astype(np.float32)
    _, per, gs, gp = po.base_train_step(s, p, pos, y)
    out = {}
    for scatter in ("segments", "atomic"):
        step = BaseStep(s.shape, p.shape, B, gpu, scatter=scatter)
        # a starting gradient of the gradient's own size (no cancellation in the check)
        g0 = to_dev(crandn(rng, H, H) * np.float32(np.sqrt(np.mean(np.abs(gs) ** 2) / 2)), gpu)
        gst, gpt = g0.clone(), torch.empty(p.shape, dtype=torch.complex64, device=gpu)
        lb = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst, gpt, accumulate=True)
        gst2 = torch.empty_like(gst)
        step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst2, gpt)
        torch.cuda.synchronize()
        e = (nrel(lb, per), nrel(gst - g0, gs), nrel(gst2, gs), nrel(gpt, gp))
        print(f"base gather {H}/{h}/B{B} spot={spot} segmented={_segmented(B, h, H)} {scatter}: loss_b {e[0]:.2e} "
              f"accumulated grad {e[1]:.2e} grad {e[2]:.2e} probe {e[3]:.2e} (vs FP64 oracle)", flush=True)
        assert max(e) < TOL
        out[scatter] = gst2
    assert nrel(out["segments"], out["atomic"]) < TOL


# rectangular objects with partial tiles along both axes: single pass (203 x 330) and segmented
# (100 x 72)
@pytest.mark.parametrize("H,W,h,B", [(203, 330, 64, 12), (100, 72, 64, 60)])
def test_base_step_gather_rectangular(gpu, H, W, h, B):
    from ptychox_amd.base import BaseStep

    rng = np.random.default_rng(H + W + B)
    s = (smooth_object(rng, max(H, W))[:H, :W] + 0.05 * crandn(rng, H, W)).astype(np.complex64)
    p = disc_probe(h, h / 4)
    pos = np.stack([rng.integers(0, W - h + 1, B), rng.integers(0, H - h + 1, B)], axis=1).astype(np.int64)
    y = np.abs(po.base_forward(s * 1.05, p, pos)["amplitude"]).astype(np.float32)
    _, per, gs, gp = po.base_train_step(s, p, pos, y)
    for scatter in ("segments", "atomic"):
        step = BaseStep(s.shape, p.shape, B, gpu, scatter=scatter)
        gst = torch.empty(s.shape, dtype=torch.complex64, device=gpu)
        gpt = torch.empty(p.shape, dtype=torch.complex64, device=gpu)
        lb = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst, gpt)
        torch.cuda.synchronize()
        e = (nrel(lb, per), nrel(gst, gs), nrel(gpt, gp))
        print(f"base gather {H}x{W}/{h}/B{B} segmented={B * h * h >= 8 * H * W} {scatter}: loss_b {e[0]:.2e} "
              f"grad {e[1]:.2e} probe {e[2]:.2e} (vs FP64 oracle)", flush=True)
        assert max(e) < TOL


@pytest.mark.parametrize("H,h,B,spot", CASES[1:3])
def test_base_segmented_gather_bitwise(gpu, H, h, B, spot):
    from ptychox_amd.base import BaseStep

    assert _segmented(B, h, H)
    rng = np.random.default_rng(5)
    s, p, pos = _base_case(rng, H, h, B, spot)
    y = to_dev(rng.uniform(0.5, 2.0, (B, h, h)).astype(np.float32), gpu)
    step = BaseStep(s.shape, p.shape, B, gpu, scatter="segments")
    grads = []
    for _ in range(3):
        gst = torch.empty((H, H), dtype=torch.complex64, device=gpu)
        gpt = torch.empty((h, h), dtype=torch.complex64, device=gpu)
        step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos, gpu), y, gst, gpt)
        # unrelated work between the calls
        _ = float(torch.randn(1 << 18, device=gpu).sum())
        grads.append(gst)
    assert torch.equal(grads[0], grads[1]) and torch.equal(grads[0], grads[2])


def _pupil(n, r, rng):
    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    rho2 = (xx**2 + yy**2) / r**2
    return ((rho2 <= 1) * np.exp(1j * rng.uniform(-1, 1) * (2 * rho2 - 1))).astype(np.complex64)


@pytest.mark.parametrize("H,h,n_side", [(64, 32, 7), (128, 64, 9)])
def test_fpm_segmented_gather_vs_oracle(gpu, H, h, n_side):
    # a dense LED grid near DC (pitch 2) with crops wrapping round the spectrum's edges
    from ptychox_amd import ptychoX_fourier, set_scatter_mode

    rng = np.random.default_rng(H + n_side)
    o = (smooth_object(rng, H, 3.0) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    q = _pupil(h, h * 0.375, rng)
    g1 = (H - h) // 2 + 2 * (np.arange(n_side) - n_side // 2)
    yy, xx = np.meshgrid(g1, g1, indexing="ij")
    led = np.stack([xx.ravel(), yy.ravel()], 1).astype(np.int64)
    led[0] = [-5, H - 3]
    led[1] = [H - 4, -7]
    B = led.shape[0]
    assert _segmented(B, h, H)
    g = rng.standard_normal((B, h, h)).astype(np.float32)
    fwd = po.fpm_forward(o, q, led)
    go, gq = po.fpm_backward(g, o.shape, q, fwd["sub_spectrum"], fwd["image_field"], led)
    got = {}
    for mode in ("segments", "atomic"):
        ot = to_dev(o, gpu).requires_grad_(True)
        qt = to_dev(q, gpu).requires_grad_(True)
        set_scatter_mode(mode)
        try:
            amp = ptychoX_fourier(ot, qt, to_dev(led, gpu), use_cuda_kernels=True)
            amp.backward(to_dev(g, gpu))
        finally:
            set_scatter_mode(None)
        e = (nrel(amp, fwd["amplitude"]), nrel(ot.grad, go), nrel(qt.grad, gq))
        print(f"fpm gather {H}/{h}/B{B} {mode}: amp {e[0]:.2e} grad_object {e[1]:.2e} grad_pupil {e[2]:.2e} "
              f"(vs FP64 oracle)", flush=True)
        assert max(e) < TOL
        got[mode] = ot.grad
    assert nrel(got["segments"], got["atomic"]) < TOL


# The forward gather passes walk a chunk of views per work-group.  With
# 1024-point rows (8 elements per thread: three LDS exchanges per transform)
# and 4 views, each work-group transforms two views, the second starting on the
# other LDS region: parity of that alternation against the oracle.
def test_base_gather_two_views_per_group_odd_exchanges(gpu):
    from ptychox_amd import ptychoX_base

    rng = np.random.default_rng(21)
    H, h, B = 1100, 1024, 4
    s = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    p = disc_probe(h, h / 4)
    pos = rng.integers(0, H - h + 1, (B, 2)).astype(np.int64)
    g = rng.standard_normal((B, h, h)).astype(np.float32)
    st, pt = to_dev(s, gpu).requires_grad_(True), to_dev(p, gpu).requires_grad_(True)
    amp = ptychoX_base(st, pt, to_dev(pos, gpu), use_cuda_kernels=True)
    amp.backward(to_dev(g, gpu))
    fwd = po.base_forward(s, p, pos)
    gs, gp = po.base_backward(g, s.shape, p, fwd["patches"], fwd["far_field"], pos)
    e = (nrel(amp, fwd["amplitude"]), nrel(st.grad, gs), nrel(pt.grad, gp))
    print(f"base {H}/{h}/B{B}: amp {e[0]:.2e} grad_sample {e[1]:.2e} grad_probe {e[2]:.2e} (vs FP64 oracle)",
          flush=True)
    assert max(e) < TOL


def test_fpm_gather_two_views_per_group_odd_exchanges(gpu):
    from ptychox_amd import ptychoX_fourier

    rng = np.random.default_rng(22)
    H, h = 1024, 1024
    o = (smooth_object(rng, H, 3.0) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    q = _pupil(h, h * 0.375, rng)
    led = np.array([[0, 0], [3, -2], [-5, 7], [1, 1]], dtype=np.int64)
    B = led.shape[0]
    g = rng.standard_normal((B, h, h)).astype(np.float32)
    ot, qt = to_dev(o, gpu).requires_grad_(True), to_dev(q, gpu).requires_grad_(True)
    amp = ptychoX_fourier(ot, qt, to_dev(led, gpu), use_cuda_kernels=True)
    amp.backward(to_dev(g, gpu))
    fwd = po.fpm_forward(o, q, led)
    go, gq = po.fpm_backward(g, o.shape, q, fwd["sub_spectrum"], fwd["image_field"], led)
    e = (nrel(amp, fwd["amplitude"]), nrel(ot.grad, go), nrel(qt.grad, gq))
    print(f"fpm {H}/{h}/B{B}: amp {e[0]:.2e} grad_object {e[1]:.2e} grad_pupil {e[2]:.2e} (vs FP64 oracle)",
          flush=True)
    assert max(e) < TOL

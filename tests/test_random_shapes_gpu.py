"""Seeded random shapes through the fused conventional and Fourier steps, in
both object-gradient modes, against the FP64 oracle (norm-wise relative error
<= 1e-5, FP32): rectangular samples and probes of independent power-of-two
sides, batches of 1 to 60 views, positions anywhere in range (conventional)
or wrapping the spectrum several periods out (Fourier)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import crandn, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5
SIDES = (8, 16, 32, 64, 128)


def _rough(rng, H, W):
    return (smooth_object(rng, max(H, W))[:H, :W] + 0.05 * crandn(rng, H, W)).astype(np.complex64)


@pytest.mark.parametrize("seed", range(8))
def test_base_step_random_shapes(gpu, seed):
    from ptychox_amd.base import BaseStep

    rng = np.random.default_rng(100 + seed)
    h, w = (int(rng.choice(SIDES)) for _ in range(2))
    H, W = int(rng.integers(h, 3 * h + 40)), int(rng.integers(w, 3 * w + 40))
    B = int(rng.integers(1, 61))
    s = _rough(rng, H, W)
    yy, xx = np.mgrid[0:h, 0:w]
    p = (np.exp(1j * rng.uniform(-1, 1)) * ((yy - h / 2) ** 2 / h**2 + (xx - w / 2) ** 2 / w**2 <= 0.2)
         + 0.01).astype(np.complex64)
    pos = np.stack([rng.integers(0, W - w + 1, B), rng.integers(0, H - h + 1, B)], axis=1).astype(np.int64)
    y = np.abs(po.base_forward(s * 1.05, p, pos)["amplitude"]).astype(np.float32)
    _, per, gs, gp = po.base_train_step(s, p, pos, y)
    for scatter in ("segments", "atomic"):
        step = BaseStep(s.shape, p.shape, B, gpu, scatter=scatter)
        gst = torch.empty(s.shape, dtype=torch.complex64, device=gpu)
        gpt = torch.empty(p.shape, dtype=torch.complex64, device=gpu)
        lb = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst, gpt)
        torch.cuda.synchronize()
        e = (nrel(lb, per), nrel(gst, gs), nrel(gpt, gp))
        print(f"base random seed {seed}: sample {H}x{W} probe {h}x{w} B{B} {scatter}: loss_b {e[0]:.2e} "
              f"grad {e[1]:.2e} probe {e[2]:.2e}", flush=True)
        assert max(e) < TOL


@pytest.mark.parametrize("seed", range(6))
def test_fpm_step_random_shapes(gpu, seed):
    from ptychox_amd.fpm import FpmStep

    rng = np.random.default_rng(200 + seed)
    h, w = (int(rng.choice(SIDES[1:])) for _ in range(2))
    H, W = h * int(rng.choice((1, 2, 4))), w * int(rng.choice((1, 2, 4)))
    B = int(rng.integers(1, 61))
    o = _rough(rng, H, W)
    q = (rng.uniform(0.5, 1.0, (h, w)) * np.exp(1j * rng.uniform(-1, 1, (h, w)))).astype(np.complex64)
    led = np.stack([rng.integers(-2 * W, 2 * W, B), rng.integers(-2 * H, 2 * H, B)], axis=1).astype(np.int64)
    y = po.fpm_forward(o * 1.1, q, led)["amplitude"].astype(np.float32)
    _, per, go, gq = po.fpm_train_step(o, q, led, y)
    for scatter in ("segments", "atomic"):
        step = FpmStep(o.shape, q.shape, B, gpu, scatter=scatter)
        got = torch.empty(o.shape, dtype=torch.complex64, device=gpu)
        gqt = torch.empty(q.shape, dtype=torch.complex64, device=gpu)
        lb = step(to_dev(o, gpu), to_dev(q, gpu), to_dev(led, gpu), to_dev(y, gpu), got, gqt)
        torch.cuda.synchronize()
        e = (nrel(lb, per), nrel(got, go), nrel(gqt, gq))
        print(f"fpm random seed {seed}: spectrum {H}x{W} pupil {h}x{w} B{B} {scatter}: loss_b {e[0]:.2e} "
              f"grad_object {e[1]:.2e} grad_pupil {e[2]:.2e}", flush=True)
        assert max(e) < TOL


@pytest.mark.parametrize("seed", range(5))
def test_nearfield_step_random_shapes(gpu, seed):
    from ptychox_amd.nearfield import NearFieldStep

    rng = np.random.default_rng(300 + seed)
    h, w = (int(rng.choice(SIDES[1:])) for _ in range(2))
    H, W = int(rng.integers(h, 3 * h + 40)), int(rng.integers(w, 3 * w + 40))
    B = int(rng.integers(1, 41))
    s = _rough(rng, H, W)
    yy, xx = np.mgrid[0:h, 0:w]
    inside = ((yy - h / 2) / (h / 2)) ** 2 + ((xx - w / 2) / (w / 2)) ** 2 <= 1
    p = (inside * np.exp(1j * rng.uniform(0, 2 * np.pi, (h, w)))).astype(np.complex64)
    fy, fx = np.fft.fftfreq(h), np.fft.fftfreq(w)
    lz = (min(h, w) / 2) ** 2 / 20.0
    t = np.exp(1j * np.pi * lz * (fy[:, None] ** 2 + fx[None, :] ** 2)).astype(np.complex64)
    pos = np.stack([rng.integers(0, W - w + 1, B), rng.integers(0, H - h + 1, B)], axis=1).astype(np.int64)
    y = po.nearfield_forward(s * 1.1, p, t, pos)["amplitude"].astype(np.float32)
    _, per, gs, gp = po.nearfield_train_step(s, p, t, pos, y)
    for scatter in ("segments", "atomic"):
        step = NearFieldStep(s.shape, p.shape, B, gpu, scatter=scatter)
        gst = torch.empty(s.shape, dtype=torch.complex64, device=gpu)
        gpt = torch.empty(p.shape, dtype=torch.complex64, device=gpu)
        lb = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(t, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst, gpt)
        torch.cuda.synchronize()
        e = (nrel(lb, per), nrel(gst, gs), nrel(gpt, gp))
        print(f"nearfield random seed {seed}: sample {H}x{W} probe {h}x{w} B{B} {scatter}: loss_b {e[0]:.2e} "
              f"grad {e[1]:.2e} probe {e[2]:.2e}", flush=True)
        assert max(e) < TOL


# coded frames of at least 32 x 32 and at least two views: smaller coded problems
# are limited by FP32 itself (the reference's torch path included)
@pytest.mark.parametrize("seed", range(5))
def test_coded_step_random_shapes(gpu, seed):
    from ptychox_amd.coded import CodedStep

    rng = np.random.default_rng(400 + seed)
    H, W = (int(rng.choice((32, 64, 128))) for _ in range(2))
    B = int(rng.integers(2, 13))
    ds = int(rng.choice((1, 2, 4)))
    o = _rough(rng, H, W)
    cells = rng.random((H // 2, W // 2)) < 0.5
    c = (np.kron(cells, np.ones((2, 2))) * np.exp(1j * np.kron(rng.uniform(0, np.pi, (H // 2, W // 2)),
                                                                 np.ones((2, 2))))).astype(np.complex64)
    fy, fx = np.fft.fftfreq(H), np.fft.fftfreq(W)
    lz = (min(H, W) / 4) ** 2 / 20.0
    t = np.exp(1j * np.pi * lz * (fy[:, None] ** 2 + fx[None, :] ** 2)).astype(np.complex64)
    sh = (rng.uniform(-4, 4, (B, 2)) / np.array([W, H])).astype(np.float32)
    y = po.coded_forward(o * 1.1, c, t, t, sh.astype(np.float64), ds)["amplitude"].astype(np.float32)
    _, per, go, gc = po.coded_train_step(o, c, t, t, sh.astype(np.float64), ds, y)
    step = CodedStep(o.shape, B, ds, gpu)
    got = torch.empty(o.shape, dtype=torch.complex64, device=gpu)
    gct = torch.empty(o.shape, dtype=torch.complex64, device=gpu)
    lb = step(to_dev(o, gpu), to_dev(c, gpu), to_dev(t, gpu), to_dev(t, gpu), to_dev(sh, gpu), to_dev(y, gpu),
              got, gct)
    torch.cuda.synchronize()
    e = (nrel(lb, per), nrel(got, go), nrel(gct, gc))
    print(f"coded random seed {seed}: frames {H}x{W} B{B} ds{ds}: loss_b {e[0]:.2e} grad_object {e[1]:.2e} "
          f"grad_coded {e[2]:.2e}", flush=True)
    assert max(e) < TOL

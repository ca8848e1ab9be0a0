"""The fused steps at the ends of the supported frame sizes, against the FP64
oracle: the largest transforms (2048 points, where the column passes take
4-column tiles and a row pass one row per work-group), the smallest (8
points), and the most elongated frames (8 x 2048, 2048 x 8).  Tolerance as
everywhere: norm-wise relative error <= 1e-5 for per-position losses and both
gradients, in both object-gradient modes where a mode applies."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import crandn, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _rough(rng, H, W):
    return (smooth_object(rng, max(H, W))[:H, :W] + 0.05 * crandn(rng, H, W)).astype(np.complex64)


def _aperture(rng, h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    return (np.exp(1j * rng.uniform(-1, 1)) * ((yy - h / 2) ** 2 / h**2 + (xx - w / 2) ** 2 / w**2 <= 0.2)
            + 0.01).astype(np.complex64)


def _transfer(h, w):
    fy, fx = np.fft.fftfreq(h), np.fft.fftfreq(w)
    lz = (min(h, w) / 2) ** 2 / 20.0
    return np.exp(1j * np.pi * lz * (fy[:, None] ** 2 + fx[None, :] ** 2)).astype(np.complex64)


def _positions(rng, H, W, h, w, B):
    return np.stack([rng.integers(0, W - w + 1, B), rng.integers(0, H - h + 1, B)], axis=1).astype(np.int64)


# (probe h, probe w, sample H, sample W, views)
SCAN = [(2048, 2048, 2048, 2048, 2), (8, 2048, 40, 2100, 5), (2048, 8, 2100, 40, 5), (8, 8, 8, 8, 3),
        (1024, 2048, 1100, 2048, 3)]


@pytest.mark.timeout(300)
@pytest.mark.parametrize("h,w,H,W,B", SCAN)
def test_base_step_extreme_frames(gpu, h, w, H, W, B):
    from ptychox_amd.base import BaseStep

    rng = np.random.default_rng(h * 3 + w + B)
    s, p = _rough(rng, H, W), _aperture(rng, h, w)
    pos = _positions(rng, H, W, h, w, B)
    y = po.base_forward(s * 1.05, p, pos)["amplitude"].astype(np.float32)
    _, per, gs, gp = po.base_train_step(s, p, pos, y)
    for scatter in ("segments", "atomic"):
        step = BaseStep(s.shape, p.shape, B, gpu, scatter=scatter)
        gst = torch.empty(s.shape, dtype=torch.complex64, device=gpu)
        gpt = torch.empty(p.shape, dtype=torch.complex64, device=gpu)
        lb = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst, gpt)
        torch.cuda.synchronize()
        e = (nrel(lb, per), nrel(gst, gs), nrel(gpt, gp))
        print(f"base sample {H}x{W} probe {h}x{w} B{B} {scatter}: loss_b {e[0]:.2e} grad {e[1]:.2e} "
              f"probe {e[2]:.2e}", flush=True)
        assert max(e) < TOL


# (pupil h, pupil w, spectrum H, spectrum W, LEDs)
FOURIER = [(2048, 2048, 2048, 2048, 2), (8, 2048, 64, 2048, 3), (2048, 8, 2048, 64, 3)]


@pytest.mark.timeout(300)
@pytest.mark.parametrize("h,w,H,W,B", FOURIER)
def test_fpm_step_extreme_frames(gpu, h, w, H, W, B):
    from ptychox_amd.fpm import FpmStep

    rng = np.random.default_rng(h + 5 * w + B)
    o = _rough(rng, H, W)
    q = (rng.uniform(0.5, 1.0, (h, w)) * np.exp(1j * rng.uniform(-1, 1, (h, w)))).astype(np.complex64)
    led = np.stack([rng.integers(-W, W, B), rng.integers(-H, H, B)], axis=1).astype(np.int64)
    y = po.fpm_forward(o * 1.1, q, led)["amplitude"].astype(np.float32)
    _, per, go, gq = po.fpm_train_step(o, q, led, y)
    for scatter in ("segments", "atomic"):
        step = FpmStep(o.shape, q.shape, B, gpu, scatter=scatter)
        got = torch.empty(o.shape, dtype=torch.complex64, device=gpu)
        gqt = torch.empty(q.shape, dtype=torch.complex64, device=gpu)
        lb = step(to_dev(o, gpu), to_dev(q, gpu), to_dev(led, gpu), to_dev(y, gpu), got, gqt)
        torch.cuda.synchronize()
        e = (nrel(lb, per), nrel(got, go), nrel(gqt, gq))
        print(f"fpm spectrum {H}x{W} pupil {h}x{w} B{B} {scatter}: loss_b {e[0]:.2e} grad_object {e[1]:.2e} "
              f"grad_pupil {e[2]:.2e}", flush=True)
        assert max(e) < TOL


@pytest.mark.timeout(300)
@pytest.mark.parametrize("h,w,H,W,B", [(2048, 2048, 2048, 2048, 2), (8, 2048, 40, 2100, 3)])
def test_nearfield_step_extreme_frames(gpu, h, w, H, W, B):
    from ptychox_amd.nearfield import NearFieldStep

    rng = np.random.default_rng(7 * h + w + B)
    s, t = _rough(rng, H, W), _transfer(h, w)
    p = (_aperture(rng, h, w) * np.exp(1j * rng.uniform(0, 2 * np.pi, (h, w)))).astype(np.complex64)
    pos = _positions(rng, H, W, h, w, B)
    y = po.nearfield_forward(s * 1.1, p, t, pos)["amplitude"].astype(np.float32)
    _, per, gs, gp = po.nearfield_train_step(s, p, t, pos, y)
    for scatter in ("segments", "atomic"):
        step = NearFieldStep(s.shape, p.shape, B, gpu, scatter=scatter)
        gst = torch.empty(s.shape, dtype=torch.complex64, device=gpu)
        gpt = torch.empty(p.shape, dtype=torch.complex64, device=gpu)
        lb = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(t, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst, gpt)
        torch.cuda.synchronize()
        e = (nrel(lb, per), nrel(gst, gs), nrel(gpt, gp))
        print(f"nearfield sample {H}x{W} probe {h}x{w} B{B} {scatter}: loss_b {e[0]:.2e} grad {e[1]:.2e} "
              f"probe {e[2]:.2e}", flush=True)
        assert max(e) < TOL


@pytest.mark.timeout(300)
@pytest.mark.parametrize("H,W,B,ds", [(2048, 2048, 2, 1), (2048, 2048, 2, 4), (64, 2048, 3, 2), (2048, 64, 3, 1)])
def test_coded_step_extreme_frames(gpu, H, W, B, ds):
    from ptychox_amd.coded import CodedStep

    rng = np.random.default_rng(H + 3 * W + B + ds)
    o = _rough(rng, H, W)
    cells = rng.random((H // 2, W // 2)) < 0.5
    c = (np.kron(cells, np.ones((2, 2))) * np.exp(1j * np.kron(rng.uniform(0, np.pi, (H // 2, W // 2)),
                                                                 np.ones((2, 2))))).astype(np.complex64)
    t = _transfer(H, W)
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
    print(f"coded frames {H}x{W} B{B} ds{ds}: loss_b {e[0]:.2e} grad_object {e[1]:.2e} grad_coded {e[2]:.2e}",
          flush=True)
    assert max(e) < TOL

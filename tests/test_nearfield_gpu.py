"""Parity of the HIP near-field path against the FP64 oracle and the unfused
torch path on the same GPU (norm-wise relative error <= 1e-5, FP32)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _diffuser(rng, n, sigma=2.0):
    from scipy.ndimage import gaussian_filter

    yy, xx = np.mgrid[0:n, 0:n] - n / 2
    disc = (xx**2 + yy**2) <= (n / 2) ** 2
    ph = gaussian_filter(rng.uniform(0, 2 * np.pi, (n, n)), sigma, mode="wrap")
    return (disc * np.exp(1j * 2 * np.pi * (ph - ph.min()) / (ph.max() - ph.min()))).astype(np.complex64)


def _transfer(n, fresnel=20.0):
    f = np.fft.fftfreq(n)
    lz = (n / 2) ** 2 / fresnel
    return np.exp(1j * np.pi * lz * (f[:, None] ** 2 + f[None, :] ** 2)).astype(np.complex64)


@pytest.mark.parametrize("H,h,B,det", [(16, 8, 3, False), (64, 32, 4, False), (256, 128, 8, False),
                                       (256, 128, 8, True), (400, 256, 5, False)])
def test_nearfield_vs_oracle(gpu, H, h, B, det):
    from ptychox_amd import ptychoX_nearfield, set_scatter_mode

    rng = np.random.default_rng(H + h + B)
    s = smooth_object(rng, H)
    p = _diffuser(rng, h)
    t = _transfer(h)
    pos = rng.integers(0, H - h + 1, (B, 2)).astype(np.int64)
    g = rng.standard_normal((B, h, h)).astype(np.float32)
    st = to_dev(s, gpu).requires_grad_(True)
    pt = to_dev(p, gpu).requires_grad_(True)
    set_scatter_mode("segments" if det else "atomic")
    try:
        amp = ptychoX_nearfield(st, pt, to_dev(t, gpu), to_dev(pos, gpu), use_cuda_kernels=True)
        amp.backward(to_dev(g, gpu))
    finally:
        set_scatter_mode(None)
    fwd = po.nearfield_forward(s, p, t, pos)
    gs, gp = po.nearfield_backward(g, s.shape, p, t, fwd["patches"], fwd["field"], pos)
    e = (nrel(amp, fwd["amplitude"]), nrel(st.grad, gs), nrel(pt.grad, gp))
    print(f"nearfield {H}/{h}/B{B} det={det}: amp {e[0]:.2e} grad_sample {e[1]:.2e} grad_probe {e[2]:.2e} "
          "(vs FP64 oracle)", flush=True)
    assert max(e) < TOL


def test_nearfield_full_size_vs_torch_path(gpu):
    """Object 1024^2, 512^2 diffuser probe, 64 positions within +-32 px, Fresnel number 20."""
    from ptychox_amd import ptychoX_nearfield

    rng = np.random.default_rng(5)
    s = smooth_object(rng, 1024, 4.0)
    p = _diffuser(rng, 512)
    t = _transfer(512)
    pos = (256 + rng.integers(-32, 33, (64, 2))).astype(np.int64)
    g = rng.standard_normal((64, 512, 512)).astype(np.float32)
    outs = []
    for native in (True, False):
        st = to_dev(s, gpu).requires_grad_(True)
        pt = to_dev(p, gpu).requires_grad_(True)
        amp = ptychoX_nearfield(st, pt, to_dev(t, gpu), to_dev(pos, gpu), use_cuda_kernels=native)
        amp.backward(to_dev(g, gpu))
        outs.append((amp.detach(), st.grad, pt.grad))
    e = [nrel(a, b) for a, b in zip(*outs)]
    print(f"nearfield 1024/512/B64: amp {e[0]:.2e} grad_sample {e[1]:.2e} grad_probe {e[2]:.2e} (vs torch path)",
          flush=True)
    assert max(e) < TOL


@pytest.mark.parametrize("det", [False, True])
def test_nearfield_step_vs_oracle(gpu, det):
    from ptychox_amd.nearfield import NearFieldStep

    rng = np.random.default_rng(9)
    s = smooth_object(rng, 200)
    p = _diffuser(rng, 64)
    t = _transfer(64)
    pos = rng.integers(0, 200 - 64 + 1, (12, 2)).astype(np.int64)
    y = po.nearfield_forward(s * 1.1, p, t, pos)["amplitude"].astype(np.float32)
    loss, per, gs, gp = po.nearfield_train_step(s, p, t, pos, y)
    step = NearFieldStep(s.shape, p.shape, 12, gpu, scatter="segments" if det else "atomic")
    gst = torch.empty(s.shape, dtype=torch.complex64, device=gpu)
    gpt = torch.empty(p.shape, dtype=torch.complex64, device=gpu)
    lb = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(t, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst, gpt)
    torch.cuda.synchronize()
    e = (nrel(lb, per), nrel(gst, gs), nrel(gpt, gp))
    print(f"nearfield step det={det}: loss_b {e[0]:.2e} grad_sample {e[1]:.2e} grad_probe {e[2]:.2e}", flush=True)
    assert max(e) < TOL


@pytest.mark.parametrize("det", [False, True])
def test_nearfield_rectangular_vs_oracle(gpu, det):
    """Sample 160 x 224 and a 64 x 32 probe with its own 64 x 32 transfer (rows x columns)."""
    from ptychox_amd import ptychoX_nearfield, set_scatter_mode

    rng = np.random.default_rng(31 + det)
    H, W, h, w, B = 160, 224, 64, 32, 6
    s = smooth_object(rng, W)[:H, :]
    yy, xx = np.mgrid[0:h, 0:w]
    inside = ((yy - h / 2) / (h / 2)) ** 2 + ((xx - w / 2) / (w / 2)) ** 2 <= 1
    p = (inside * np.exp(1j * rng.uniform(0, 2 * np.pi, (h, w)))).astype(np.complex64)
    fy, fx = np.fft.fftfreq(h), np.fft.fftfreq(w)
    t = np.exp(1j * np.pi * 60.0 * (fy[:, None] ** 2 + fx[None, :] ** 2)).astype(np.complex64)
    pos = np.stack([rng.integers(0, W - w + 1, B), rng.integers(0, H - h + 1, B)], axis=1).astype(np.int64)
    g = rng.standard_normal((B, h, w)).astype(np.float32)
    st = to_dev(s, gpu).requires_grad_(True)
    pt = to_dev(p, gpu).requires_grad_(True)
    set_scatter_mode("segments" if det else "atomic")
    try:
        amp = ptychoX_nearfield(st, pt, to_dev(t, gpu), to_dev(pos, gpu), use_cuda_kernels=True)
        amp.backward(to_dev(g, gpu))
    finally:
        set_scatter_mode(None)
    fwd = po.nearfield_forward(s, p, t, pos)
    gs, gp = po.nearfield_backward(g, s.shape, p, t, fwd["patches"], fwd["field"], pos)
    e = (nrel(amp, fwd["amplitude"]), nrel(st.grad, gs), nrel(pt.grad, gp))
    print(f"nearfield 160x224/64x32/B6 det={det}: amp {e[0]:.2e} grad_sample {e[1]:.2e} grad_probe {e[2]:.2e} "
          "(vs FP64 oracle)", flush=True)
    assert max(e) < TOL

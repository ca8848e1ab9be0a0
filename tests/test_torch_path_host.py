"""The drop-in operators' torch path (``use_cuda_kernels=False``, the branch
the reference runs without its extension: python/ptychoX/base/ops.py:60-105,
fpm/ops.py:56-97, coded/ops.py:45-53,136-171) against the FP64 oracle, on the
CPU: forward amplitudes, and the gradients autograd returns for the drivers'
MSE loss (python/train_ptychography.py:555), norm-wise relative error <= 1e-5
in FP32.  The GPU tests at the benchmark sizes compare the HIP kernels with
this same path, so it is pinned here independently."""

from __future__ import annotations

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import ptycho_oracle as po
from tests.helpers import crandn, disc_probe, nrel, smooth_object

TOL = 1e-5


def _grads(amp_fn, leaves, measured):
    leaves = [torch.from_numpy(x).requires_grad_(True) for x in leaves]
    amp = amp_fn(*leaves)
    loss = F.mse_loss(amp, torch.from_numpy(measured))
    loss.backward()
    return amp.detach(), float(loss.detach()), [t.grad for t in leaves]


@pytest.mark.parametrize("H,h,B", [(96, 32, 7), (160, 64, 12)])
def test_base_torch_path_vs_oracle(H, h, B):
    from ptychox_amd import ptychoX_base

    rng = np.random.default_rng(H + B)
    s = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    p = disc_probe(h, h / 4)
    pos = rng.integers(0, H - h + 1, (B, 2)).astype(np.int64)
    y = (po.base_forward(s * 1.1, p, pos)["amplitude"]).astype(np.float32)
    loss, _, gs_ref, gp_ref = po.base_train_step(s, p, pos, y)
    amp, l32, (gs, gp) = _grads(lambda a, b: ptychoX_base(a, b, torch.from_numpy(pos)), [s, p], y)
    assert nrel(amp, po.base_forward(s, p, pos)["amplitude"]) < TOL
    assert abs(l32 - loss) / loss < TOL
    assert nrel(gs, gs_ref) < TOL
    assert nrel(gp, gp_ref) < TOL


@pytest.mark.parametrize("H,h,B", [(64, 16, 9), (128, 32, 25)])
def test_fpm_torch_path_vs_oracle(H, h, B):
    from ptychox_amd import ptychoX_fourier

    rng = np.random.default_rng(3 * H + B)
    o = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    q = disc_probe(h, h / 3, defocus=1e9)
    led = rng.integers(-H // 2, H // 2 - h, (B, 2)).astype(np.int64)
    y = po.fpm_forward(o * 0.9, q, led)["amplitude"].astype(np.float32)
    loss, _, go_ref, gq_ref = po.fpm_train_step(o, q, led, y)
    amp, l32, (go, gq) = _grads(lambda a, b: ptychoX_fourier(a, b, torch.from_numpy(led)), [o, q], y)
    assert nrel(amp, po.fpm_forward(o, q, led)["amplitude"]) < TOL
    assert abs(l32 - loss) / loss < TOL
    assert nrel(go, go_ref) < TOL
    assert nrel(gq, gq_ref) < TOL


def test_nearfield_torch_path_vs_oracle():
    from ptychox_amd import angular_spectrum_transfer, ptychoX_nearfield

    rng = np.random.default_rng(21)
    H, h, B = 128, 64, 10
    s = (smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64)
    p = (disc_probe(h, h / 2) * np.exp(1j * rng.uniform(0, 2 * np.pi, (h, h)))).astype(np.complex64)
    t = angular_spectrum_transfer(h, h, (h / 2) ** 2 / 20.0).numpy()
    pos = rng.integers(0, H - h + 1, (B, 2)).astype(np.int64)
    y = po.nearfield_forward(s * 1.1, p, t, pos)["amplitude"].astype(np.float32)
    loss, _, gs_ref, gp_ref = po.nearfield_train_step(s, p, t, pos, y)
    amp, l32, (gs, gp) = _grads(lambda a, b: ptychoX_nearfield(a, b, torch.from_numpy(t), torch.from_numpy(pos)),
                                [s, p], y)
    assert nrel(amp, po.nearfield_forward(s, p, t, pos)["amplitude"]) < TOL
    assert abs(l32 - loss) / loss < TOL
    assert nrel(gs, gs_ref) < TOL
    assert nrel(gp, gp_ref) < TOL


@pytest.mark.parametrize("ds", [1, 2])
def test_coded_torch_path_vs_oracle(ds):
    from ptychox_amd import angular_spectrum_transfer, ptychoX_coded

    rng = np.random.default_rng(30 + ds)
    n, B = 64, 9
    o = (smooth_object(rng, n) + 0.05 * crandn(rng, n, n)).astype(np.complex64)
    c = (rng.integers(0, 2, (n, n)) * np.exp(1j * rng.uniform(0, np.pi, (n, n)))).astype(np.complex64)
    t = angular_spectrum_transfer(n, n, (n / 2) ** 2 / 20.0).numpy()
    sh = (rng.integers(-3, 4, (B, 2)) / n).astype(np.float32)
    y = po.coded_forward(o * 1.1, c, t, t, sh, ds)["amplitude"].astype(np.float32)
    loss, _, go_ref, gc_ref = po.coded_train_step(o, c, t, t, sh, ds, y)
    tt, st = torch.from_numpy(t), torch.from_numpy(sh)
    amp, l32, (go, gc) = _grads(lambda a, b: ptychoX_coded(a, b, tt, tt, st, ds), [o, c], y)
    assert nrel(amp, po.coded_forward(o, c, t, t, sh, ds)["amplitude"]) < TOL
    assert abs(l32 - loss) / loss < TOL
    assert nrel(go, go_ref) < TOL
    assert nrel(gc, gc_ref) < TOL

"""Parity of the HIP conventional-ptychography path (C ABI via ctypes) against
the FP64 oracle and the unfused torch path on the same GPU.

Tolerance: norm-wise relative error <= 1e-5 in FP32 for amplitudes, per-position
losses and each gradient; the FFT alone <= 2e-6.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import crandn, disc_probe, nrel, raster_positions, smooth_object, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _lib():
    from ptychox_amd import _lib

    _lib.lib()
    return _lib


@pytest.mark.parametrize("nbytes", [16, 4096 * 16 + 48, 64 << 20])
def test_stream_copy_is_exact(gpu, nbytes):
    lib = _lib()
    src = torch.randint(-(2**31), 2**31 - 1, (nbytes // 4,), dtype=torch.int32, device=gpu)
    dst = torch.full((nbytes // 4 + 4,), 7, dtype=torch.int32, device=gpu)  # 16 guard bytes past the end
    lib.call("ptx_stream_copy", lib.ptr(src), lib.ptr(dst), nbytes, lib.stream_of(gpu))
    torch.cuda.synchronize()
    assert torch.equal(dst[: nbytes // 4], src)
    assert torch.all(dst[nbytes // 4 :] == 7)


@pytest.mark.parametrize("h,w", [(8, 8), (16, 64), (64, 64), (128, 32), (256, 256), (512, 512), (1024, 128), (2048, 16)])
@pytest.mark.parametrize("inverse", [0, 1])
def test_fft2d_matches_torch(gpu, h, w, inverse):
    lib = _lib()
    rng = np.random.default_rng(h * 7 + w + inverse)
    x = to_dev(crandn(rng, 3, h, w), gpu)
    ref = torch.fft.ifft2(x) * (h * w) if inverse else torch.fft.fft2(x)
    y = x.clone()
    lib.call("ptx_fft2d", lib.ptr(y), 3, h, w, inverse, lib.stream_of(gpu))
    torch.cuda.synchronize()
    print(f"fft2d {h}x{w} inverse={inverse}: rel {nrel(y, ref):.2e}", flush=True)
    assert nrel(y, ref) < 2e-6


def _case(rng, H, h, B, device):
    s = smooth_object(rng, H)
    p = disc_probe(h, h / 4)
    pos = np.stack([rng.integers(0, H - h + 1, B), rng.integers(0, H - h + 1, B)], axis=1).astype(np.int64)
    return s, p, pos


@pytest.mark.parametrize("H,h,B", [(96, 32, 5), (256, 64, 7), (300, 128, 9), (512, 256, 3)])
def test_base_forward_backward_vs_oracle(gpu, H, h, B):
    from ptychox_amd import ptychoX_base

    rng = np.random.default_rng(H + h + B)
    s, p, pos = _case(rng, H, h, B, gpu)
    g = rng.standard_normal((B, h, h)).astype(np.float32)
    st = to_dev(s, gpu).requires_grad_(True)
    pt = to_dev(p, gpu).requires_grad_(True)
    amp = ptychoX_base(st, pt, to_dev(pos, gpu), use_cuda_kernels=True)
    amp.backward(to_dev(g, gpu))
    fwd = po.base_forward(s, p, pos)
    gs, gp = po.base_backward(g, s.shape, p, fwd["patches"], fwd["far_field"], pos)
    e = (nrel(amp, fwd["amplitude"]), nrel(st.grad, gs), nrel(pt.grad, gp))
    print(f"base {H}/{h}/B{B}: amp {e[0]:.2e} grad_sample {e[1]:.2e} grad_probe {e[2]:.2e} (vs FP64 oracle)", flush=True)
    assert max(e) < TOL


@pytest.mark.parametrize("det", [False, True])
def test_base_native_vs_torch_path_full_size(gpu, det):
    """The headline workload's shapes: object 2048^2, probe 512^2, B=64."""
    from ptychox_amd import ptychoX_base, set_scatter_mode

    rng = np.random.default_rng(2)
    s = smooth_object(rng, 2048, 4.0)
    p = disc_probe(512, 128)
    pos = raster_positions(rng, 30, 52, 14, 3)[:64]
    g = rng.standard_normal((64, 512, 512)).astype(np.float32)
    set_scatter_mode("segments" if det else "atomic")
    try:
        outs = []
        for native in (True, False):
            st = to_dev(s, gpu).requires_grad_(True)
            pt = to_dev(p, gpu).requires_grad_(True)
            amp = ptychoX_base(st, pt, to_dev(pos, gpu), use_cuda_kernels=native)
            amp.backward(to_dev(g, gpu))
            outs.append((amp.detach(), st.grad, pt.grad))
    finally:
        set_scatter_mode(None)
    e = [nrel(a, b) for a, b in zip(outs[0], outs[1])]
    print(f"base 2048/512/B64 det={det}: amp {e[0]:.2e} grad_sample {e[1]:.2e} grad_probe {e[2]:.2e} (vs torch path)", flush=True)
    assert max(e) < TOL


@pytest.mark.parametrize("scatter", ["segments", "atomic"])
def test_base_fused_step_full_size_vs_torch_path(gpu, scatter):
    """The fused step at the headline shapes (object 2048^2, probe 512^2, the
    first 64 positions of the 30 x 30 raster: a band of the object, as the
    benchmark's mini-batches) against the reference torch path with
    F.mse_loss: per-position losses and both gradients."""
    from ptychox_amd import ptychoX_base
    from ptychox_amd.base import BaseStep

    rng = np.random.default_rng(6)
    s = to_dev(smooth_object(rng, 2048, 4.0), gpu)
    p = to_dev(disc_probe(512, 128), gpu)
    pos = to_dev(raster_positions(rng, 30, 52, 14, 3)[:64], gpu)
    with torch.no_grad():
        y = ptychoX_base(s * 1.05, p, pos, use_cuda_kernels=False).contiguous()
    st, pt = s.clone().requires_grad_(True), p.clone().requires_grad_(True)
    amp = ptychoX_base(st, pt, pos, use_cuda_kernels=False)
    torch.nn.functional.mse_loss(amp, y).backward()
    per = ((amp.detach() - y) ** 2).mean(dim=(1, 2))
    step = BaseStep(s.shape, p.shape, 64, gpu, scatter=scatter)
    gs, gp = torch.empty_like(s), torch.empty_like(p)
    lb = step(s, p, pos, y, gs, gp)
    torch.cuda.synchronize()
    e = (nrel(lb, per), nrel(gs, st.grad), nrel(gp, pt.grad))
    print(f"base fused step 2048/512/B64 {scatter}: loss_b {e[0]:.2e} grad_sample {e[1]:.2e} "
          f"grad_probe {e[2]:.2e} (vs torch path)", flush=True)
    assert max(e) < TOL


def test_base_deterministic_bitwise(gpu):
    from ptychox_amd import ptychoX_base, set_scatter_mode

    rng = np.random.default_rng(3)
    s, p, pos = _case(rng, 400, 128, 24, gpu)
    g = to_dev(rng.standard_normal((24, 128, 128)).astype(np.float32), gpu)
    set_scatter_mode("segments")
    try:
        grads = []
        for _ in range(2):
            st = to_dev(s, gpu).requires_grad_(True)
            pt = to_dev(p, gpu).requires_grad_(True)
            ptychoX_base(st, pt, to_dev(pos, gpu), use_cuda_kernels=True).backward(g)
            grads.append((st.grad.clone(), pt.grad.clone()))
    finally:
        set_scatter_mode(None)
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])


@pytest.mark.parametrize("det", [False, True])
def test_base_step_vs_oracle(gpu, det):
    from ptychox_amd.base import BaseStep

    rng = np.random.default_rng(4)
    s, p, pos = _case(rng, 320, 128, 10, gpu)
    y = np.abs(po.base_forward(s * 1.05, p, pos)["amplitude"]).astype(np.float32)
    loss, per, gs, gp = po.base_train_step(s, p, pos, y)
    step = BaseStep(s.shape, p.shape, 10, gpu, scatter="segments" if det else "atomic", keep_amplitude=True)
    gst = torch.empty(s.shape, dtype=torch.complex64, device=gpu)
    gpt = torch.empty(p.shape, dtype=torch.complex64, device=gpu)
    lb = step(to_dev(s, gpu), to_dev(p, gpu), to_dev(pos, gpu), to_dev(y, gpu), gst, gpt)
    torch.cuda.synchronize()
    e = (nrel(lb, per), nrel(gst, gs), nrel(gpt, gp))
    print(f"base step det={det}: loss_b {e[0]:.2e} grad_sample {e[1]:.2e} grad_probe {e[2]:.2e} (vs FP64 oracle)", flush=True)
    assert max(e) < TOL


def test_base_rejects_out_of_range(gpu):
    from ptychox_amd import ptychoX_base

    s = torch.zeros(64, 64, dtype=torch.complex64, device=gpu)
    p = torch.zeros(16, 16, dtype=torch.complex64, device=gpu)
    with pytest.raises(ValueError):
        ptychoX_base(s, p, torch.tensor([[50, 0]], device=gpu), use_cuda_kernels=True)

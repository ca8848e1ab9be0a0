"""The drop-in boundary on the GPU.

* The ``_C`` replacement modules (``ptychox_amd.ref_c``), driven exactly as
  the reference's autograd Functions drive their extension
  (python/ptychoX/base/ops.py:52-54,82-91, fpm/ops.py:51-53,73-83,
  coded/ops.py:86-96,118-130): forward, keep what it returns, backward with
  those tensors.  Results must equal the package's own autograd operators
  bitwise (the same kernels run).
* Recompute instead of store: gradients are bitwise the same whether the
  forward's per-pattern outputs were saved or recomputed in the backward.
* A native operator's forward and backward captured in a CUDA graph (no host
  synchronisation inside), replayed with eager work in between.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.helpers import crandn, disc_probe, smooth_object, to_dev

pytestmark = pytest.mark.gpu


def _base_inputs(gpu, seed=1, H=256, h=64, B=12):
    rng = np.random.default_rng(seed)
    s = to_dev((smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64), gpu)
    p = to_dev(disc_probe(h, h / 4), gpu)
    pos = to_dev(rng.integers(0, H - h + 1, (B, 2)).astype(np.int64), gpu)
    g = to_dev(rng.uniform(-1, 1, (B, h, h)).astype(np.float32), gpu)
    return s, p, pos, g


def _fpm_inputs(gpu, seed=2, H=256, h=64, B=25):
    rng = np.random.default_rng(seed)
    o = to_dev((smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64), gpu)
    q = to_dev(disc_probe(h, h / 3, defocus=1e9), gpu)
    led = to_dev(rng.integers(-H // 2, H // 2 - h, (B, 2)).astype(np.int64), gpu)
    g = to_dev(rng.uniform(-1, 1, (B, h, h)).astype(np.float32), gpu)
    return o, q, led, g


def _coded_inputs(gpu, seed=3, n=128, B=9, ds=2):
    from ptychox_amd import angular_spectrum_transfer

    rng = np.random.default_rng(seed)
    o = to_dev((smooth_object(rng, n) + 0.05 * crandn(rng, n, n)).astype(np.complex64), gpu)
    c = to_dev((rng.integers(0, 2, (n, n)) * np.exp(1j * rng.uniform(0, np.pi, (n, n)))).astype(np.complex64), gpu)
    t = angular_spectrum_transfer(n, n, (n / 2) ** 2 / 20.0, device=gpu)
    sh = to_dev((rng.integers(-3, 4, (B, 2)) / n).astype(np.float32), gpu)
    g = to_dev(rng.uniform(-1, 1, (B, n // ds, n // ds)).astype(np.float32), gpu)
    return o, c, t, sh, ds, g


def _autograd(fn, leaves, g, *rest):
    leaves = [x.clone().requires_grad_(True) for x in leaves]
    amp = fn(*leaves, *rest)
    grads = torch.autograd.grad(amp, leaves, grad_outputs=g)
    torch.cuda.synchronize()
    return amp.detach(), grads


def test_base_c_module_matches_the_package(gpu):
    from ptychox_amd import ptychoX_base
    from ptychox_amd.ref_c import base as C

    s, p, pos, g = _base_inputs(gpu)
    patch_stack, far_field, amp = C.forward(s, p, pos)
    gs, gp = C.backward(g, patch_stack, p, far_field, pos, *s.shape)
    torch.cuda.synchronize()
    amp_ref, (gs_ref, gp_ref) = _autograd(lambda a, b: ptychoX_base(a, b, pos, use_cuda_kernels=True), [s, p], g)
    assert torch.equal(amp, amp_ref)
    assert torch.equal(gs, gs_ref) and torch.equal(gp, gp_ref)
    with pytest.raises(RuntimeError, match="probe must be complex64"):
        C.forward(s, p.real.contiguous(), pos)
    with pytest.raises(RuntimeError, match="larger than sample"):
        C.forward(s[:32, :32].contiguous(), p, pos)


def test_fpm_c_module_matches_the_package(gpu):
    from ptychox_amd import ptychoX_fourier
    from ptychox_amd.ref_c import fpm as C

    o, q, led, g = _fpm_inputs(gpu)
    sub, image_field, amp = C.forward(o, q, led)
    go, gq = C.backward(g, sub, q, image_field, led, *o.shape)
    torch.cuda.synchronize()
    amp_ref, (go_ref, gq_ref) = _autograd(lambda a, b: ptychoX_fourier(a, b, led, use_cuda_kernels=True), [o, q], g)
    assert torch.equal(amp, amp_ref)
    assert torch.equal(go, go_ref) and torch.equal(gq, gq_ref)
    with pytest.raises(RuntimeError, match="led_indices_xy must be int64"):
        C.forward(o, q, led.to(torch.int32))


def test_coded_c_module_matches_the_package(gpu):
    from ptychox_amd import ptychoX_coded
    from ptychox_amd.ref_c import coded as C

    o, c, t, sh, ds, g = _coded_inputs(gpu)
    spec, x_fwd, sensor, amp = C.forward(o, c, t, t, sh, ds)
    go, gc = C.backward(g, spec, x_fwd, sensor, c, t, t, sh, ds)
    torch.cuda.synchronize()
    amp_ref, (go_ref, gc_ref) = _autograd(
        lambda a, b: ptychoX_coded(a, b, t, t, sh, ds, use_cuda_kernels=True), [o, c], g)
    assert torch.equal(amp, amp_ref)
    assert torch.equal(go, go_ref) and torch.equal(gc, gc_ref)
    with pytest.raises(RuntimeError, match="shifts_xy must be float32"):
        C.forward(o, c, t, t, sh.double(), ds)


def _recompute_pair(fn, leaves, g, *rest):
    from ptychox_amd import set_recompute

    out = {}
    try:
        for mode in ("never", "always"):
            set_recompute(mode)
            out[mode] = _autograd(fn, leaves, g, *rest)
    finally:
        set_recompute(None)
    return out["never"], out["always"]


def test_recompute_gives_the_stored_gradients(gpu):
    from ptychox_amd import angular_spectrum_transfer, ptychoX_base, ptychoX_coded, ptychoX_fourier, ptychoX_nearfield

    s, p, pos, g = _base_inputs(gpu)
    cases = [("base", lambda a, b: ptychoX_base(a, b, pos, use_cuda_kernels=True), [s, p], g)]
    t = angular_spectrum_transfer(64, 64, 32.0 ** 2 / 20.0, device=gpu)
    cases.append(("nearfield", lambda a, b: ptychoX_nearfield(a, b, t, pos, use_cuda_kernels=True), [s, p], g))
    o, q, led, gf = _fpm_inputs(gpu)
    cases.append(("fourier", lambda a, b: ptychoX_fourier(a, b, led, use_cuda_kernels=True), [o, q], gf))
    oc, c, tc, sh, ds, gc = _coded_inputs(gpu)
    cases.append(("coded", lambda a, b: ptychoX_coded(a, b, tc, tc, sh, ds, use_cuda_kernels=True), [oc, c], gc))
    for name, fn, leaves, grad in cases:
        (amp0, g0), (amp1, g1) = _recompute_pair(fn, leaves, grad)
        assert torch.equal(amp0, amp1), name
        for a, b in zip(g0, g1):
            assert torch.equal(a, b), name


def test_recompute_saves_only_inputs(gpu):
    from ptychox_amd import ptychoX_coded, set_recompute

    oc, c, tc, sh, ds, _ = _coded_inputs(gpu)
    try:
        for mode, want_stacks in (("never", 2), ("always", 0)):
            set_recompute(mode)
            a = oc.clone().requires_grad_(True)
            amp = ptychoX_coded(a, c, tc, tc, sh, ds, use_cuda_kernels=True)
            saved = [x for x in amp.grad_fn.saved_tensors if x is not None]
            stacks = [x for x in saved if x.dim() == 3]  # [batch, H, W] per-translation tensors
            assert len(stacks) == want_stacks, mode
    finally:
        set_recompute(None)


def test_base_operator_in_a_captured_graph(gpu):
    from ptychox_amd import ptychoX_base

    s, p, pos, g = _base_inputs(gpu, seed=5)
    sl, pl = s.clone().requires_grad_(True), p.clone().requires_grad_(True)

    def run():
        amp = ptychoX_base(sl, pl, pos, use_cuda_kernels=True)
        gs, gp = torch.autograd.grad(amp, (sl, pl), grad_outputs=g)
        return amp, gs, gp

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        eager = [x.clone() for x in run()]  # warm-up off the capture stream
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    # fresh leaves for the capture (the warm-up's autograd nodes belong to the side stream)
    sl, pl = s.clone().requires_grad_(True), p.clone().requires_grad_(True)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        static = run()
    for k in range(3):
        x = torch.randn(1 << 20, device=gpu)  # eager work between the replays
        (x * 2).sum().item()
        graph.replay()
        torch.cuda.synchronize()
        for a, b in zip(static, eager):
            assert torch.equal(a, b), k


@pytest.mark.parametrize("bad", [(-1, 5), (5, -1), (193, 5), (5, 193)])
def test_gpu_positions_outside_the_sample_raise(gpu, bad):
    """On the HIP path, positions already on the GPU are checked beside the
    forward kernels (their reduction runs on a side stream and is waited for
    after the kernels are queued); the operators still raise the host check's
    error, and in-range batches give the same amplitudes as CPU positions,
    which are checked before anything is launched."""
    from ptychox_amd import angular_spectrum_transfer, ptychoX_base, ptychoX_nearfield

    s, p, pos, _ = _base_inputs(gpu)  # 256^2 sample, 64^2 probe: corners in [0, 192]
    t = angular_spectrum_transfer(64, 64, 40.0, device=gpu)
    for op, extra in ((ptychoX_base, ()), (ptychoX_nearfield, (t,))):
        amp_gpu = op(s, p, *extra, pos, use_cuda_kernels=True)
        amp_cpu = op(s, p, *extra, pos.cpu(), use_cuda_kernels=True)
        torch.cuda.synchronize()
        assert torch.equal(amp_gpu, amp_cpu)
        outside = pos.clone()
        outside[3] = torch.tensor(bad, device=gpu)
        with pytest.raises(ValueError, match="scan window out of range"):
            op(s, p, *extra, outside, use_cuda_kernels=True)
        with pytest.raises(ValueError, match="scan window out of range"):
            op(s, p, *extra, outside.cpu(), use_cuda_kernels=True)
        # the stream is usable afterwards and the next in-range call is unchanged
        assert torch.equal(op(s, p, *extra, pos, use_cuda_kernels=True), amp_gpu)

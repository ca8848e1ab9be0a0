"""Empty batches through every fused step and operator: no error, gradient
outputs zero-filled (left unchanged when accumulating), empty loss and
amplitude outputs.  The reference's torch path cannot stack an empty batch
(python/ptychoX/base/ops.py:60-66); the C ABI defines the case (a step or
adjoint of no views contributes nothing), and the Python layer passes it on."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.helpers import crandn, to_dev

pytestmark = pytest.mark.gpu


def _garbage(shape, gpu):
    return torch.full(shape, complex(3.0, -2.0), dtype=torch.complex64, device=gpu)


def _check_step(step_call, g1, g2):
    # plain: both gradients become zero; accumulating: both stay as they were
    loss = step_call(accumulate=False)
    torch.cuda.synchronize()
    assert loss.numel() == 0
    assert torch.count_nonzero(g1) == 0 and torch.count_nonzero(g2) == 0
    g1.fill_(complex(1.0, 2.0))
    g2.fill_(complex(-1.0, 0.5))
    a1, a2 = g1.clone(), g2.clone()
    step_call(accumulate=True)
    torch.cuda.synchronize()
    assert torch.equal(g1, a1) and torch.equal(g2, a2)


def test_steps_with_no_views(gpu):
    from ptychox_amd.base import BaseStep
    from ptychox_amd.coded import CodedStep
    from ptychox_amd.fpm import FpmStep
    from ptychox_amd.nearfield import NearFieldStep

    rng = np.random.default_rng(0)
    s = to_dev(crandn(rng, 128, 128), gpu)
    p = to_dev(crandn(rng, 32, 32), gpu)
    pos = torch.zeros((4, 2), dtype=torch.int64, device=gpu)
    y = torch.ones((4, 32, 32), dtype=torch.float32, device=gpu)
    t = to_dev(crandn(rng, 32, 32), gpu)
    for name, step, call in (
        ("conventional", BaseStep(s.shape, p.shape, 4, gpu), lambda st, g1, g2, acc: st(s, p, pos, y, g1, g2, batch=0,
                                                                                          accumulate=acc)),
        ("conventional, atomic", BaseStep(s.shape, p.shape, 4, gpu, scatter="atomic"),
         lambda st, g1, g2, acc: st(s, p, pos, y, g1, g2, batch=0, accumulate=acc)),
        ("fourier", FpmStep(s.shape, p.shape, 4, gpu), lambda st, g1, g2, acc: st(s, p, pos, y, g1, g2, batch=0,
                                                                                     accumulate=acc)),
        ("near-field", NearFieldStep(s.shape, p.shape, 4, gpu),
         lambda st, g1, g2, acc: st(s, p, t, pos, y, g1, g2, batch=0, accumulate=acc)),
    ):
        g1, g2 = _garbage(s.shape, gpu), _garbage(p.shape, gpu)
        print(name, flush=True)
        _check_step(lambda accumulate: call(step, g1, g2, accumulate), g1, g2)
    o = to_dev(crandn(rng, 64, 64), gpu)
    c = to_dev(crandn(rng, 64, 64), gpu)
    tc = to_dev(crandn(rng, 64, 64), gpu)
    sh = torch.zeros((3, 2), dtype=torch.float32, device=gpu)
    yc = torch.ones((3, 64, 64), dtype=torch.float32, device=gpu)
    for loss in ("mse", "fd-isotropic"):
        step = CodedStep(o.shape, 3, 1, gpu, loss=loss)
        g1, g2 = _garbage(o.shape, gpu), _garbage(c.shape, gpu)
        print("coded", loss, flush=True)
        _check_step(lambda accumulate: step(o, c, tc, tc, sh, yc, g1, g2, batch=0, accumulate=accumulate), g1, g2)


def test_operators_with_no_positions(gpu):
    from ptychox_amd import ptychoX_base, ptychoX_coded, ptychoX_fourier, ptychoX_nearfield

    rng = np.random.default_rng(1)
    s = to_dev(crandn(rng, 128, 128), gpu).requires_grad_(True)
    p = to_dev(crandn(rng, 32, 32), gpu).requires_grad_(True)
    t = to_dev(crandn(rng, 32, 32), gpu)
    none = torch.zeros((0, 2), dtype=torch.int64, device=gpu)
    for name, op in (("conventional", lambda: ptychoX_base(s, p, none, use_cuda_kernels=True)),
                     ("fourier", lambda: ptychoX_fourier(s, p, none, use_cuda_kernels=True)),
                     ("near-field", lambda: ptychoX_nearfield(s, p, t, none, use_cuda_kernels=True))):
        s.grad = p.grad = None
        amp = op()
        assert amp.shape == (0, 32, 32), (name, amp.shape)
        amp.sum().backward()
        torch.cuda.synchronize()
        assert s.grad is not None and torch.count_nonzero(s.grad) == 0, name
        assert p.grad is not None and torch.count_nonzero(p.grad) == 0, name
    o = to_dev(crandn(rng, 64, 64), gpu).requires_grad_(True)
    c = to_dev(crandn(rng, 64, 64), gpu).requires_grad_(True)
    tc = to_dev(crandn(rng, 64, 64), gpu)
    amp = ptychoX_coded(o, c, tc, tc, torch.zeros((0, 2), dtype=torch.float32, device=gpu), 2, use_cuda_kernels=True)
    assert amp.shape == (0, 32, 32)
    amp.sum().backward()
    torch.cuda.synchronize()
    assert torch.count_nonzero(o.grad) == 0 and torch.count_nonzero(c.grad) == 0

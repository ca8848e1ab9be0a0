"""FP32 torch-CPU restatement of the reference's torch fallback path.

TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline; never shipped).

It keeps the reference's op order *including* its per-view Python loops, so it
is the "reference CPU path" that the benchmark times on the host cores:

* conventional forward  -- ``python/ptychoX/base/ops.py:60-66`` (per-view slice
  loop + ``torch.stack``, ``fft2``, ``fftshift``, ``abs``);
* conventional backward -- ``base/ops.py:93-105`` (phase clamp, ``ifft2`` of the
  ``ifftshift`` times ``h*w``, per-view ``+=`` loop);
* FPM forward/backward  -- ``python/ptychoX/fpm/ops.py:56-59`` and ``85-97``
  (flat-index gather, ``index_add_``);
* coded                  -- ``python/ptychoX/coded/ops.py:136-154`` under autograd;
* near-field             -- crop form of the MATLAB near-field operator
  (``matlab/PtychoX_updated/scripts/forward/fwd_adPtychoX_nearf.cu:58-114``).
"""

from __future__ import annotations

import math

import torch
import torch.nn.functional as F

EPS = torch.finfo(torch.float32).eps


def base_forward(sample: torch.Tensor, probe: torch.Tensor, positions_xy: torch.Tensor):
    h, w = probe.shape
    patches = []
    for x, y in positions_xy.tolist():
        patches.append(sample[y : y + h, x : x + w])
    patch_stack = torch.stack(patches, dim=0)
    far_field = torch.fft.fftshift(torch.fft.fft2(patch_stack * probe.unsqueeze(0)), dim=(-2, -1))
    return far_field.abs(), patch_stack, far_field


def base_backward(grad_output, sample_shape, probe, patch_stack, far_field, positions_xy):
    h, w = probe.shape
    phase = far_field / far_field.abs().clamp_min(EPS)
    grad_exit = torch.fft.ifft2(torch.fft.ifftshift(grad_output * phase, dim=(-2, -1))) * (h * w)
    grad_sample = torch.zeros(sample_shape, dtype=probe.dtype)
    grad_probe = torch.zeros_like(probe)
    pconj = probe.conj()
    for view in range(patch_stack.shape[0]):
        x = int(positions_xy[view, 0])
        y = int(positions_xy[view, 1])
        grad_sample[y : y + h, x : x + w] += grad_exit[view] * pconj
        grad_probe += grad_exit[view] * patch_stack[view].conj()
    return grad_sample, grad_probe


def base_step(sample, probe, positions_xy, measured):
    """One conventional mini-batch fwd + MSE + adjoint, the timed CPU unit."""
    amp, patches, far = base_forward(sample, probe, positions_xy)
    diff = amp - measured
    loss = (diff * diff).mean()
    g = 2.0 * diff / diff.numel()
    gs, gp = base_backward(g, tuple(sample.shape), probe, patches, far, positions_xy)
    return loss, gs, gp


def _fpm_idx(oshape, lshape, led):
    oh, ow = oshape
    lh, lw = lshape
    yy = (led[:, 1, None, None] + torch.arange(lh)[None, :, None] + oh // 2) % oh
    xx = (led[:, 0, None, None] + torch.arange(lw)[None, None, :] + ow // 2) % ow
    return yy * ow + xx


def fpm_forward(object_field, pupil, led):
    offt = torch.fft.fft2(object_field)
    idx = _fpm_idx(object_field.shape, pupil.shape, led)
    sub = offt.reshape(-1).gather(0, idx.reshape(-1)).reshape(led.shape[0], *pupil.shape)
    field = torch.fft.ifft2(torch.fft.fftshift(sub * pupil.unsqueeze(0), dim=(-2, -1)))
    return field.abs(), sub, field


def fpm_backward(grad_output, object_shape, pupil, sub, field, led):
    lh, lw = pupil.shape
    oh, ow = object_shape
    phase = field / field.abs().clamp_min(EPS)
    glow = torch.fft.ifftshift(torch.fft.fft2(grad_output * phase) / (lh * lw), dim=(-2, -1))
    gpupil = torch.sum(glow * sub.conj(), dim=0)
    idx = _fpm_idx(object_shape, pupil.shape, led)
    gflat = torch.zeros(oh * ow, dtype=pupil.dtype)
    gflat.index_add_(0, idx.reshape(-1), (glow * pupil.conj().unsqueeze(0)).reshape(-1))
    return torch.fft.ifft2(gflat.reshape(oh, ow)) * (oh * ow), gpupil


def coded_forward(object_field, coded_surface, t1, t2, shifts_xy, ds):
    h, w = object_field.shape
    fy = torch.fft.fftfreq(h).to(torch.float32) * h
    fx = torch.fft.fftfreq(w).to(torch.float32) * w
    yy, xx = torch.meshgrid(fy, fx, indexing="ij")
    arg = 2.0 * math.pi * (shifts_xy[:, 0, None, None] * xx[None] + shifts_xy[:, 1, None, None] * yy[None])
    ramp = torch.exp(1j * arg)
    spec = torch.fft.fft2(object_field)
    x_fwd = torch.fft.ifft2(spec[None] * t1[None] * ramp)
    sensor = torch.fft.ifft2(torch.fft.fft2(x_fwd * coded_surface[None]) * t2[None])
    pooled = F.avg_pool2d(sensor.abs().square().unsqueeze(1), kernel_size=ds, stride=ds)
    return torch.sqrt(pooled[:, 0].clamp_min(0.0))


def nearfield_forward(sample, probe, transfer, positions_xy):
    h, w = probe.shape
    patches = []
    for x, y in positions_xy.tolist():
        patches.append(sample[y : y + h, x : x + w])
    stack = torch.stack(patches, dim=0)
    return torch.fft.ifft2(transfer[None] * torch.fft.fft2(stack * probe[None])).abs()


def base_epoch(sample, probe, positions_xy, measured, batch, order):
    """fwd + MSE + adjoint over every position in mini-batches of ``batch`` in
    the given order (``python/train_ptychography.py:531-553``: ``randperm``,
    slices of the permutation, a partial last batch); returns the gradients of
    the last mini-batch."""
    out = None
    for start in range(0, positions_xy.shape[0], batch):
        bi = order[start : start + batch]
        out = base_step(sample, probe, positions_xy[bi], measured[bi])
    return out


def base_train(sample, probe, positions_xy, measured, batch, epochs, lr, seed=0):
    """The reference driver's reconstruction loop on the CPU
    (``python/train_ptychography.py:473-603``): per epoch a ``randperm`` of the
    positions, mini-batches through the literal forward / MSE / adjoint, one
    ``torch.optim.Adam`` step (betas (0.9, 0.999), eps 1e-16) per mini-batch on
    object and probe, cosine learning rate (460-465) set at each epoch start.
    Returns the final object, probe and the last epoch's mean loss."""
    s = torch.nn.Parameter(sample.clone())
    p = torch.nn.Parameter(probe.clone())
    opt = torch.optim.Adam([s, p], lr=lr, betas=(0.9, 0.999), eps=1e-16)
    gen = torch.Generator().manual_seed(seed)
    P = positions_xy.shape[0]
    last = 0.0
    for epoch in range(1, epochs + 1):
        phase = 1.0 if epochs <= 1 else (1.0 + math.cos(math.pi * (epoch - 1) / epochs)) * 0.5
        for g in opt.param_groups:
            g["lr"] = lr * phase
        order = torch.randperm(P, generator=gen)
        total = 0.0
        for start in range(0, P, batch):
            bi = order[start : start + batch]
            with torch.no_grad():
                loss, gs, gp = base_step(s.detach(), p.detach(), positions_xy[bi], measured[bi])
            s.grad, p.grad = gs, gp
            opt.step()
            total += float(loss) * len(bi)
        last = total / P
    return s.detach(), p.detach(), last

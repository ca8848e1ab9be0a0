"""Device-resident reconstruction loop with hipGraph epoch capture.

The caller of the operators in the reference is a plain PyTorch loop
(``python/train_ptychography.py:473-603``): per epoch a ``torch.randperm`` of the
scan positions, mini-batches of it, ``F.mse_loss`` against the measured
amplitudes, ``loss.backward()`` and ``torch.optim.Adam(betas=(0.9, 0.999),
eps=1e-16)`` on the complex object and probe (two parameter groups, cosine
learning-rate decay per epoch, ``cosine_lrs`` at 460-465).

Here the same epoch runs as one captured graph: for each mini-batch the fused
HIP step (``ptx_base_step_indexed``: forward, amplitude loss, adjoint, reading
positions and measurements through the permutation without copies) followed by
``ptx_adam_step`` on the object and on the probe.  Everything the graph reads
that changes between epochs -- the permutation and the Adam step sizes -- lives
in device buffers rewritten before each replay; per-position losses stay on the
device and are reduced once per epoch.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import _lib
from ._utils import require_native_frame, scatter_flags


def cosine_lr(base_lr: float, min_lr: float, epochs: int, epoch: int, scale: float = 1.0) -> float:
    """``cosine_lrs`` of python/train_ptychography.py:460-465 for one group (epoch is 1-based)."""
    phase = 1.0 if epochs <= 1 else (1.0 + math.cos(math.pi * (epoch - 1) / epochs)) * 0.5
    return (min_lr + (base_lr - min_lr) * phase) * scale


def adam_hyper(lr: float, beta1: float, beta2: float, eps: float, step: int) -> tuple[float, float, float]:
    """``(lr / (1 - beta1^t), sqrt(1 - beta2^t), eps)`` as torch's foreach Adam forms them."""
    bc1 = 1.0 - beta1**step
    bc2 = 1.0 - beta2**step
    return lr / bc1, bc2**0.5, eps


class AdamState:
    """First and second moments of one complex parameter (per real component).

    ``real_nonneg=True`` follows every update with the coded driver's
    constraint on the coded surface, ``p = max(Re p, 0) + 0j``
    (python/train_coded.py:240-241), in the same kernel."""

    def __init__(self, param: torch.Tensor, beta1: float = 0.9, beta2: float = 0.999, real_nonneg: bool = False):
        self.param = param
        self.m = torch.zeros_like(torch.view_as_real(param))
        self.v = torch.zeros_like(self.m)
        self.beta1, self.beta2 = float(beta1), float(beta2)
        self.flags = _lib.PTX_ADAM_PROJECT_REAL_NONNEG if real_nonneg else 0

    def step(self, grad: torch.Tensor, hyper: torch.Tensor) -> None:
        """One update; ``hyper`` is a device float[3] (see :func:`adam_hyper`)."""
        _lib.call(
            "ptx_adam_step",
            _lib.ptr(self.param), _lib.ptr(grad), _lib.ptr(self.m), _lib.ptr(self.v), self.m.numel(),
            self.beta1, self.beta2, _lib.ptr(hyper), self.flags, _lib.stream_of(self.param.device),
        )


@dataclass
class EpochResult:
    epoch: int
    loss: float
    lrs: tuple[float, float]


class ConventionalReconstruction:
    """Conventional-ptychography reconstruction with the reference driver's
    semantics (mini-batches of a fresh permutation each epoch, MSE amplitude loss
    per batch, Adam on object and probe, cosine learning rate), on the HIP path.

    ``graph=True`` captures one epoch the first time it runs and replays it.
    """

    def __init__(self, sample: torch.Tensor, probe: torch.Tensor, positions: torch.Tensor, measured: torch.Tensor,
                 batch_size: int, lr: float, probe_lr: float | None = None, min_lr: float = 0.0, epochs: int = 1,
                 generator: torch.Generator | None = None, graph: bool = True, scatter: str | None = None,
                 betas=(0.9, 0.999), eps: float = 1e-16):
        dev = sample.device
        if dev.type != "cuda":
            raise ValueError("the reconstruction loop runs on the GPU")
        self.device = dev
        self.sample = sample.detach().to(torch.complex64).contiguous().clone()
        self.probe = probe.detach().to(torch.complex64).contiguous().clone()
        self.positions = positions.to(device=dev, dtype=torch.int64).contiguous()
        self.measured = measured.to(device=dev, dtype=torch.float32).contiguous()
        self.P = int(self.positions.shape[0])
        self.H, self.W = self.sample.shape
        self.h, self.w = self.probe.shape
        require_native_frame(self.h, self.w, "probe")
        if self.measured.shape != (self.P, self.h, self.w):
            raise ValueError("measured must have shape [positions, probe_h, probe_w]")
        self.batch = int(batch_size)
        self.nb = (self.P + self.batch - 1) // self.batch
        self.base_lrs = (float(lr), float(probe_lr if probe_lr is not None else lr))
        self.min_lr, self.epochs = float(min_lr), int(epochs)
        self.betas, self.eps = (float(betas[0]), float(betas[1])), float(eps)
        self.generator = generator
        self.flags = scatter_flags(scatter)
        self.graph_enabled = bool(graph)
        self.graph: torch.cuda.CUDAGraph | None = None
        self.steps_done = 0
        # device state
        self.gs = torch.empty_like(self.sample)
        self.gp = torch.empty_like(self.probe)
        self.opt_s = AdamState(self.sample, *self.betas)
        self.opt_p = AdamState(self.probe, *self.betas)
        self.order = torch.arange(self.P, device=dev, dtype=torch.int64)
        self.loss = torch.zeros(self.P, dtype=torch.float32, device=dev)
        self.hyper = torch.zeros((2, self.nb, 4), dtype=torch.float32, device=dev)
        self.hyper_host = torch.zeros((2, self.nb, 4), dtype=torch.float32, pin_memory=True)
        nbytes = _lib.query("ptx_base_step_workspace", self.H, self.W, self.h, self.w, self.batch, self.flags)
        self.ws = _lib.workspace(nbytes, dev)

    # one epoch of launches, all on the current stream (captured when graphed)
    def _launch_epoch(self) -> None:
        stream = _lib.stream_of(self.device)
        for k in range(self.nb):
            start = k * self.batch
            B = min(self.batch, self.P - start)
            _lib.call(
                "ptx_base_step_indexed",
                _lib.ptr(self.sample), self.H, self.W, _lib.ptr(self.probe), self.h, self.w,
                _lib.ptr(self.positions), _lib.ptr(self.measured), self.P,
                ctypes_offset(self.order, start), B, 1.0 / (B * self.h * self.w),
                ctypes_offset(self.loss, start), None, _lib.ptr(self.gs), _lib.ptr(self.gp),
                _lib.ptr(self.ws), self.ws.numel(), self.flags, stream,
            )
            self.opt_s.step(self.gs, self.hyper[0, k])
            self.opt_p.step(self.gp, self.hyper[1, k])

    def _prepare(self, epoch: int, scale: float = 1.0) -> tuple[float, float]:
        if self.generator is not None:
            perm = torch.randperm(self.P, generator=self.generator, device=self.device)
        else:
            perm = torch.randperm(self.P, device=self.device)
        self.order.copy_(perm)
        lrs = tuple(cosine_lr(b, self.min_lr, self.epochs, epoch, scale) for b in self.base_lrs)
        rows = [[list(adam_hyper(lr, self.betas[0], self.betas[1], self.eps, self.steps_done + k + 1)) + [0.0]
                 for k in range(self.nb)] for lr in lrs]
        torch.cuda.current_stream(self.device).synchronize()  # the previous copy out of the pinned buffer is done
        self.hyper_host.copy_(torch.tensor(rows, dtype=torch.float32))
        self.hyper.copy_(self.hyper_host, non_blocking=True)
        return lrs

    def run_epoch(self, epoch: int, scale: float = 1.0) -> EpochResult:
        lrs = self._prepare(epoch, scale)
        if not self.graph_enabled:
            self._launch_epoch()
        else:
            if self.graph is None:
                self._capture()
            self.graph.replay()
        self.steps_done += self.nb
        epoch_loss = float(self.loss.sum()) / self.P  # one sync per epoch
        return EpochResult(epoch, epoch_loss, lrs)

    def _capture(self) -> None:
        # The first epoch runs inside the capture's warm-up on a side stream,
        # then a second, captured copy is recorded; the caller's state must
        # therefore be snapshotted and restored around the warm-up.
        snap = [t.clone() for t in (self.sample, self.probe, self.opt_s.m, self.opt_s.v, self.opt_p.m, self.opt_p.v)]
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._launch_epoch()
        torch.cuda.current_stream(self.device).wait_stream(side)
        for t, s in zip((self.sample, self.probe, self.opt_s.m, self.opt_s.v, self.opt_p.m, self.opt_p.v), snap):
            t.copy_(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch_epoch()
        self.graph = g

    def fit(self, epochs: int | None = None) -> list[EpochResult]:
        return [self.run_epoch(e) for e in range(1, (epochs or self.epochs) + 1)]


def ctypes_offset(t: torch.Tensor, index: int):
    """Device pointer to element ``index`` of a contiguous 1-D tensor."""
    import ctypes

    return ctypes.c_void_p(t.data_ptr() + index * t.element_size())

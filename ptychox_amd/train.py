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
``ptx_adam_step`` on the object and on the probe.  The Fourier, near-field and
coded drivers (python/train_fpm.py:206-283, python/train_coded.py:182-262) run
the same way with their fused steps after an on-device gather of the
mini-batch's LEDs / positions / shifts and measurements; the coded one adds
the fd losses and the real non-negative coded-surface constraint.  Everything the graph reads
that changes between epochs -- the permutation and the Adam step sizes -- lives
in device buffers rewritten before each replay; per-position losses stay on the
device and are reduced once per epoch.  Nothing in an epoch waits for the host:
the step sizes upload from alternating page-locked buffers, and the epoch loss
is reduced and copied behind the epoch and read only when asked for, so epochs
are launched back to back.
"""

from __future__ import annotations

import math

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

    def native(self, hyper: torch.Tensor) -> "_lib.PtxAdam":
        """This state as the ``ptx_adam`` a fused training step takes
        (no projection: the fused steps update unconstrained unknowns)."""
        if self.flags:
            raise ValueError("the fused Adam update has no projection")
        return _lib.PtxAdam(self.m.data_ptr(), self.v.data_ptr(), hyper.data_ptr(), self.beta1, self.beta2)

    def step(self, grad: torch.Tensor, hyper: torch.Tensor) -> None:
        """One update; ``hyper`` is a device float[3] (see :func:`adam_hyper`)."""
        _lib.call(
            "ptx_adam_step",
            _lib.ptr(self.param), _lib.ptr(grad), _lib.ptr(self.m), _lib.ptr(self.v), self.m.numel(),
            self.beta1, self.beta2, _lib.ptr(hyper), self.flags, _lib.stream_of(self.param.device),
        )


class EpochResult:
    """One epoch: its number, the learning rates of the two parameter groups
    and the drivers' epoch loss.  The loss of an epoch run on the device is
    reduced behind the epoch and copied to page-locked host memory; reading
    ``loss`` waits for that copy only (the first time), so the loop itself never
    synchronises with the device."""

    __slots__ = ("epoch", "lrs", "_value", "_host", "_event", "_count")

    def __init__(self, epoch: int, loss: float | None, lrs: tuple[float, float], *, host=None, event=None,
                 count: int = 1):
        self.epoch = int(epoch)
        self.lrs = lrs
        self._value = None if loss is None else float(loss)
        self._host, self._event, self._count = host, event, count

    @property
    def loss(self) -> float:
        if self._value is None:
            self._event.synchronize()
            self._value = float(self._host) / self._count
            self._host = self._event = None
        return self._value

    def __repr__(self) -> str:
        return f"EpochResult(epoch={self.epoch}, loss={self.loss!r}, lrs={self.lrs!r})"


class _EpochLoop:
    """Epoch machinery shared by the reconstructions: a fresh permutation each
    epoch, mini-batches of it, one fused step per mini-batch into the gradient
    buffers, Adam on the object and the auxiliary unknown (probe / pupil /
    coded surface) with per-group cosine learning rates, per-view losses on the
    device, and the whole epoch captured as a graph (``graph=True``) and
    replayed.  Subclasses implement ``_batch``."""

    def __init__(self, obj: torch.Tensor, aux: torch.Tensor, count: int, batch_size: int, lr: float,
                 aux_lr: float | None, min_lr: float, epochs: int, generator: torch.Generator | None, graph: bool,
                 betas, eps: float, aux_real_nonneg: bool = False, train_aux: bool = True):
        dev = obj.device
        if dev.type != "cuda":
            raise ValueError("the reconstruction loop runs on the GPU")
        self.device = dev
        self.obj = obj.detach().to(torch.complex64).contiguous().clone()
        self.aux = aux.detach().to(device=dev, dtype=torch.complex64).contiguous().clone()
        self.P = int(count)
        self.batch = int(batch_size)
        if self.batch < 1 or self.P < 1:
            raise ValueError("batch_size and the number of views must be positive")
        self.nb = (self.P + self.batch - 1) // self.batch
        self.base_lrs = (float(lr), float(aux_lr if aux_lr is not None else lr))
        self.min_lr, self.epochs = float(min_lr), int(epochs)
        self.betas, self.eps = (float(betas[0]), float(betas[1])), float(eps)
        self.generator = generator
        self.train_aux = bool(train_aux)
        self.graph_enabled = bool(graph)
        self.graph: torch.cuda.CUDAGraph | None = None
        self.steps_done = 0
        self.g_obj = torch.empty_like(self.obj)
        self.g_aux = torch.empty_like(self.aux)
        self.opt_o = AdamState(self.obj, *self.betas)
        self.opt_a = AdamState(self.aux, *self.betas, real_nonneg=aux_real_nonneg)
        self.order = torch.arange(self.P, device=dev, dtype=torch.int64)
        self.loss = torch.zeros(self.P, dtype=torch.float32, device=dev)
        self.hyper = torch.zeros((2, self.nb, 4), dtype=torch.float32, device=dev)
        # two page-locked staging buffers, alternating by epoch: a buffer is
        # rewritten only once its previous upload (two epochs back) has finished
        self.hyper_host = [torch.zeros((2, self.nb, 4), dtype=torch.float32, pin_memory=True) for _ in range(2)]
        self._hyper_done: list[torch.cuda.Event | None] = [None, None]
        self._prepared = 0

    def _batch(self, k: int, start: int, B: int) -> None:
        """Launch mini-batch k (views order[start:start+B]) on the current
        stream: gradients into g_obj / g_aux, per-view losses into
        loss[start:start+B]."""
        raise NotImplementedError

    def _launch_epoch(self) -> None:
        for k in range(self.nb):
            start = k * self.batch
            B = min(self.batch, self.P - start)
            self._batch(k, start, B)
            self.opt_o.step(self.g_obj, self.hyper[0, k])
            if self.train_aux:
                self.opt_a.step(self.g_aux, self.hyper[1, k])

    def _prepare(self, epoch: int, scale: float = 1.0) -> tuple[float, float]:
        if self.generator is not None:
            perm = torch.randperm(self.P, generator=self.generator, device=self.device)
        else:
            perm = torch.randperm(self.P, device=self.device)
        self.order.copy_(perm)
        lrs = tuple(cosine_lr(b, self.min_lr, self.epochs, epoch, scale) for b in self.base_lrs)
        rows = [[list(adam_hyper(lr, self.betas[0], self.betas[1], self.eps, self.steps_done + k + 1)) + [0.0]
                 for k in range(self.nb)] for lr in lrs]
        j = self._prepared % 2
        self._prepared += 1
        if self._hyper_done[j] is not None:
            self._hyper_done[j].synchronize()  # this buffer's upload two epochs ago has finished
        stream = torch.cuda.current_stream(self.device)
        self.hyper_host[j].copy_(torch.tensor(rows, dtype=torch.float32))
        self.hyper.copy_(self.hyper_host[j], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(stream)
        self._hyper_done[j] = ev
        return lrs

    def _epoch_result(self, epoch: int, lrs) -> EpochResult:
        # the drivers' epoch_loss: sum over mini-batches of loss * batch, / views,
        # reduced on the device behind the epoch and read when asked for
        stream = torch.cuda.current_stream(self.device)
        host = torch.empty((), dtype=torch.float32, pin_memory=True)
        host.copy_(self.loss.sum(), non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(stream)
        return EpochResult(epoch, None, lrs, host=host, event=ev, count=self.P)

    def run_epoch(self, epoch: int, scale: float = 1.0) -> EpochResult:
        lrs = self._prepare(epoch, scale)
        if not self.graph_enabled:
            self._launch_epoch()
        else:
            if self.graph is None:
                self._capture()
            self.graph.replay()
        self.steps_done += self.nb
        return self._epoch_result(epoch, lrs)

    def _state(self):
        return (self.obj, self.aux, self.opt_o.m, self.opt_o.v, self.opt_a.m, self.opt_a.v)

    def _capture(self) -> None:
        # The first epoch runs inside the capture's warm-up on a side stream,
        # then a second, captured copy is recorded; the caller's state must
        # therefore be snapshotted and restored around the warm-up.
        snap = [t.clone() for t in self._state()]
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._launch_epoch()
        torch.cuda.current_stream(self.device).wait_stream(side)
        for t, s in zip(self._state(), snap):
            t.copy_(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch_epoch()
        self.graph = g

    def fit(self, epochs: int | None = None) -> list[EpochResult]:
        return [self.run_epoch(e) for e in range(1, (epochs or self.epochs) + 1)]


class ConventionalReconstruction(_EpochLoop):
    """Conventional-ptychography reconstruction with the reference driver's
    semantics (mini-batches of a fresh permutation each epoch, MSE amplitude loss
    per batch, Adam on object and probe, cosine learning rate), on the HIP path.
    The fused step reads positions and measurements through the permutation
    (``ptx_base_step_indexed``), so no mini-batch is gathered.

    ``graph=True`` captures one epoch the first time it runs and replays it.
    ``fused_adam=True`` (the default) runs each mini-batch as one
    ``ptx_base_train_step``: the Adam updates of object and probe are applied
    in the epilogues that finish the gradients, with the same results as the
    step followed by two ``ptx_adam_step`` calls (``fused_adam=False``).
    """

    def __init__(self, sample: torch.Tensor, probe: torch.Tensor, positions: torch.Tensor, measured: torch.Tensor,
                 batch_size: int, lr: float, probe_lr: float | None = None, min_lr: float = 0.0, epochs: int = 1,
                 generator: torch.Generator | None = None, graph: bool = True, scatter: str | None = None,
                 betas=(0.9, 0.999), eps: float = 1e-16, fused_adam: bool = True):
        dev = sample.device
        positions = positions.to(device=dev, dtype=torch.int64).contiguous()
        super().__init__(sample, probe, positions.shape[0], batch_size, lr, probe_lr, min_lr, epochs, generator,
                         graph, betas, eps)
        self.positions = positions
        self.measured = measured.to(device=dev, dtype=torch.float32).contiguous()
        self.H, self.W = self.obj.shape
        self.h, self.w = self.aux.shape
        require_native_frame(self.h, self.w, "probe")
        if self.measured.shape != (self.P, self.h, self.w):
            raise ValueError("measured must have shape [positions, probe_h, probe_w]")
        self.flags = scatter_flags(scatter)
        self.fused_adam = bool(fused_adam)
        nbytes = _lib.query("ptx_base_step_workspace", self.H, self.W, self.h, self.w, self.batch, self.flags)
        self.ws = _lib.workspace(nbytes, dev)

    @property
    def sample(self) -> torch.Tensor:
        return self.obj

    @property
    def probe(self) -> torch.Tensor:
        return self.aux

    def _launch_epoch(self) -> None:
        if not self.fused_adam:
            super()._launch_epoch()
            return
        import ctypes

        for k in range(self.nb):
            start = k * self.batch
            B = min(self.batch, self.P - start)
            ao = self.opt_o.native(self.hyper[0, k])
            ap = self.opt_a.native(self.hyper[1, k]) if self.train_aux else None
            _lib.call(
                "ptx_base_train_step",
                _lib.ptr(self.obj), self.H, self.W, _lib.ptr(self.aux), self.h, self.w,
                _lib.ptr(self.positions), _lib.ptr(self.measured), self.P,
                ctypes_offset(self.order, start), B, 1.0 / (B * self.h * self.w),
                ctypes_offset(self.loss, start), None, _lib.ptr(self.g_obj), _lib.ptr(self.g_aux),
                ctypes.c_void_p(ctypes.addressof(ao)), None if ap is None else ctypes.c_void_p(ctypes.addressof(ap)),
                _lib.ptr(self.ws), self.ws.numel(), self.flags, _lib.stream_of(self.device),
            )

    def _batch(self, k: int, start: int, B: int) -> None:
        _lib.call(
            "ptx_base_step_indexed",
            _lib.ptr(self.obj), self.H, self.W, _lib.ptr(self.aux), self.h, self.w,
            _lib.ptr(self.positions), _lib.ptr(self.measured), self.P,
            ctypes_offset(self.order, start), B, 1.0 / (B * self.h * self.w),
            ctypes_offset(self.loss, start), None, _lib.ptr(self.g_obj), _lib.ptr(self.g_aux),
            _lib.ptr(self.ws), self.ws.numel(), self.flags, _lib.stream_of(self.device),
        )


class _GatheredReconstruction(_EpochLoop):
    """Mini-batches gathered on the device (``index_select`` of the views'
    geometry and measurements into fixed buffers, inside the captured graph),
    then one fused step of ``self.step``."""

    def _init_gather(self, geometry: torch.Tensor, measured: torch.Tensor) -> None:
        dev = self.device
        self.geometry = geometry.contiguous()
        self.measured = measured.to(device=dev, dtype=torch.float32).contiguous()
        if self.geometry.shape[0] != self.P or self.measured.shape[0] != self.P:
            raise ValueError("geometry and measured must have one entry per view")
        self.geo_buf = torch.empty((self.batch,) + tuple(self.geometry.shape[1:]), dtype=self.geometry.dtype,
                                   device=dev)
        self.meas_buf = torch.empty((self.batch,) + tuple(self.measured.shape[1:]), dtype=torch.float32, device=dev)

    def _gather(self, start: int, B: int):
        idx = self.order[start:start + B]
        torch.index_select(self.geometry, 0, idx, out=self.geo_buf[:B])
        torch.index_select(self.measured, 0, idx, out=self.meas_buf[:B])
        return self.geo_buf[:B], self.meas_buf[:B]


class FourierReconstruction(_GatheredReconstruction):
    """Fourier-ptychography reconstruction with python/train_fpm.py:206-283's
    semantics: object and pupil, LED mini-batches of a fresh permutation each
    epoch, ``F.mse_loss``, Adam (eps 1e-16) with cosine learning rates."""

    def __init__(self, object_field, pupil, led_indices_xy, measured, batch_size, lr, pupil_lr=None, min_lr=0.0,
                 epochs=1, generator=None, graph=True, scatter=None, betas=(0.9, 0.999), eps=1e-16):
        from .fpm import FpmStep

        dev = object_field.device
        led = led_indices_xy.to(device=dev, dtype=torch.int64).contiguous()
        super().__init__(object_field, pupil, led.shape[0], batch_size, lr, pupil_lr, min_lr, epochs, generator,
                         graph, betas, eps)
        self._init_gather(led, measured)
        self.step = FpmStep(self.obj.shape, self.aux.shape, self.batch, dev, scatter=scatter)

    def _batch(self, k: int, start: int, B: int) -> None:
        led, meas = self._gather(start, B)
        self.step(self.obj, self.aux, led, meas, self.g_obj, self.g_aux, batch=B)
        self.loss[start:start + B].copy_(self.step.loss[:B])


class NearFieldReconstruction(_GatheredReconstruction):
    """Near-field reconstruction (object and probe, fixed transfer function)
    with the conventional driver's loop semantics."""

    def __init__(self, sample, probe, transfer, positions, measured, batch_size, lr, probe_lr=None, min_lr=0.0,
                 epochs=1, generator=None, graph=True, scatter=None, betas=(0.9, 0.999), eps=1e-16):
        from .nearfield import NearFieldStep

        dev = sample.device
        pos = positions.to(device=dev, dtype=torch.int64).contiguous()
        super().__init__(sample, probe, pos.shape[0], batch_size, lr, probe_lr, min_lr, epochs, generator, graph,
                         betas, eps)
        self._init_gather(pos, measured)
        self.transfer = transfer.to(device=dev, dtype=torch.complex64).contiguous()
        self.step = NearFieldStep(self.obj.shape, self.aux.shape, self.batch, dev, scatter=scatter)

    def _batch(self, k: int, start: int, B: int) -> None:
        pos, meas = self._gather(start, B)
        self.step(self.obj, self.aux, self.transfer, pos, meas, self.g_obj, self.g_aux, batch=B)
        self.loss[start:start + B].copy_(self.step.loss[:B])


class CodedReconstruction(_GatheredReconstruction):
    """Coded-ptychography reconstruction with python/train_coded.py:182-262's
    semantics: object and coded surface, shift mini-batches of a fresh
    permutation each epoch, ``--loss`` mse / fd-isotropic / fd-anisotropic,
    Adam (eps 1e-16) with cosine learning rates, the coded surface constrained
    to real non-negative values after every step (``real_coded_surface``,
    default on) or kept fixed (``fix_coded_surface``)."""

    def __init__(self, object_field, coded_surface, transfer_1, transfer_2, shifts_xy, downsample, measured,
                 batch_size, lr, coded_lr=None, min_lr=0.0, epochs=1, generator=None, graph=True, loss="mse",
                 real_coded_surface=True, fix_coded_surface=False, betas=(0.9, 0.999), eps=1e-16):
        from .coded import CodedStep

        dev = object_field.device
        sh = shifts_xy.to(device=dev, dtype=torch.float32).contiguous()
        super().__init__(object_field, coded_surface, sh.shape[0], batch_size, lr, coded_lr, min_lr, epochs,
                         generator, graph, betas, eps,
                         aux_real_nonneg=bool(real_coded_surface) and not fix_coded_surface,
                         train_aux=not fix_coded_surface)
        self._init_gather(sh, measured)
        self.t1 = None if transfer_1 is None else transfer_1.to(device=dev, dtype=torch.complex64).contiguous()
        self.t2 = transfer_2.to(device=dev, dtype=torch.complex64).contiguous()
        self.ds = int(downsample)
        self.loss_name = loss
        self.step = CodedStep(self.obj.shape, self.batch, self.ds, dev, loss=loss)

    def _batch(self, k: int, start: int, B: int) -> None:
        sh, meas = self._gather(start, B)
        self.step(self.obj, self.aux, self.t1, self.t2, sh, meas, self.g_obj, self.g_aux, batch=B)
        self.loss[start:start + B].copy_(self.step.loss[:B])

    def _epoch_result(self, epoch: int, lrs) -> EpochResult:
        if self.loss_name == "mse":
            return super()._epoch_result(epoch, lrs)
        # fd losses are sums over a batch: epoch_loss = sum_k B_k * (sum over batch k) / views
        # (read at once: this path synchronises once per epoch)
        per = self.loss.cpu()
        tot = 0.0
        for k in range(self.nb):
            s0 = k * self.batch
            Bk = min(self.batch, self.P - s0)
            tot += Bk * float(per[s0:s0 + Bk].sum())
        return EpochResult(epoch, tot / self.P, lrs)


def ctypes_offset(t: torch.Tensor, index: int):
    """Device pointer to element ``index`` of a contiguous 1-D tensor."""
    import ctypes

    return ctypes.c_void_p(t.data_ptr() + index * t.element_size())

"""Continuous-batching reconstruction server on one GPU.

A request is one reconstruction -- conventional, Fourier, near-field or coded
ptychography, of any supported size -- run for a fixed number of Adam epochs
(one full-data fused step of the request's operator, the amplitude MSE loss of
the reference drivers, then Adam on both unknowns; python/train_ptychography.py
:473-603 is the per-request loop this replaces).  The server keeps every
admitted request's unknowns, Adam moments, measurements and workspace resident
in HBM and runs the requests concurrently, each on a stream of its own:

* admission by HBM footprint: a queued request is admitted when a slot is free
  and its bytes -- computed from its shapes with the library's workspace
  queries, before anything is allocated -- fit the budget beside the requests
  already resident;
* continuous batching: every pass of the loop retires finished requests (a
  non-blocking event query), frees their memory and admits arrivals at once,
  so requests join and leave while others are mid-reconstruction;
* each request has at most ``depth`` epochs queued on the device; the loop
  tops every request up in turn, so long and short requests interleave even
  where several streams share one hardware queue;
* the first epoch of a request runs eagerly, the second is captured as a graph
  on the request's stream and every later epoch replays it (the Adam step
  sizes are read from a device table indexed by an epoch counter the graph
  itself advances, so nothing is rewritten between replays).

A request's arithmetic touches only its own buffers and uses no atomics
(sorted-segment gradients, fixed-order reductions), so its result is bitwise
the same alone and in any mix.
"""

from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .train import AdamState, adam_hyper, cosine_lr

KINDS = ("conventional", "fourier", "nearfield", "coded")


@dataclass(order=True)
class Request:
    arrival: float                      # seconds after the start of the run
    rid: int = field(compare=True)
    kind: str = field(compare=False, default="conventional")
    size: int = field(compare=False, default=128)    # probe / pupil / frame side
    views: int = field(compare=False, default=100)   # scan positions, LEDs or translations
    seed: int = field(compare=False, default=0)
    epochs: int = field(compare=False, default=20)
    lr: float = field(compare=False, default=2e-2)


# (kind, frame side, object side, views range)
REQUEST_CLASSES = (
    ("conventional", 128, 512, (100, 400)),
    ("conventional", 256, 1024, (100, 400)),
    ("fourier", 128, 512, (81, 225)),
    ("nearfield", 512, 1024, (64, 64)),
    ("coded", 512, 512, (25, 25)),
)
_OBJECT_SIDE = {(k, n): o for k, n, o, _ in REQUEST_CLASSES}


def standard_mix(n: int, rate: float, seed: int = 0, epochs: int = 20, classes=REQUEST_CLASSES,
                 first_id: int = 0) -> list[Request]:
    """``n`` requests with Poisson arrivals at ``rate`` per second (``inf``: all
    at t = 0), uniform over the request classes: conventional probe 128 on a
    512^2 object and probe 256 on 1024^2, each with U{100..400} positions;
    Fourier 128^2 pupil on a 512^2 object with 81 or 225 LEDs; near-field
    probe 512 on 1024^2 with 64 positions; coded 512^2 with 25 translations.
    Each request runs ``epochs`` Adam epochs; its data seed is its id."""
    rng = np.random.default_rng(seed)
    t = np.zeros(n) if math.isinf(rate) else np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        kind, size, _obj, (lo, hi) = classes[int(rng.integers(0, len(classes)))]
        if kind == "fourier":
            views = int(rng.choice([lo, hi]))
        else:
            views = int(rng.integers(lo, hi + 1))
        rid = first_id + i
        reqs.append(Request(float(t[i]), rid, kind, size, views, seed=rid, epochs=epochs))
    return reqs


# ----------------------------------------------------------------------------- geometry and data

def object_side(req: Request) -> int:
    """Object side of a request: the class table's, else 4x the frame
    (conventional, Fourier), 2x (near-field) or 1x (coded)."""
    o = _OBJECT_SIDE.get((req.kind, req.size))
    if o is not None:
        return o
    return req.size * {"conventional": 4, "fourier": 4, "nearfield": 2, "coded": 1}[req.kind]


def _conventional_geometry(n: int, obj: int, views: int) -> tuple[int, int, int, int]:
    """(scan side, step, origin, jitter) of a square raster of ``views``
    positions that covers the ``obj`` x ``obj`` object with an ``n`` x ``n``
    probe (positions stay in range with the jitter)."""
    side = int(math.ceil(math.sqrt(views)))
    origin, jitter = 8, 2
    step = 1 if side == 1 else max((obj - n - origin - jitter) // (side - 1), 1)
    if origin - jitter < 0 or origin + step * (side - 1) + jitter + n > obj:
        raise ValueError(f"{views} positions of a {n}^2 probe do not fit a {obj}^2 object")
    return side, step, origin, jitter


def shapes(req: Request) -> dict:
    """Object, auxiliary (probe / pupil / mask) and frame shapes of a request."""
    if req.kind not in KINDS:
        raise ValueError(f"unknown request kind {req.kind!r}")
    o, n = object_side(req), req.size
    return {"object": (o, o), "aux": (n, n), "frame": (n, n), "views": req.views}


@dataclass
class Problem:
    """Host-side data of one request (what arrives with it): the initial
    guesses and the measurements."""

    kind: str
    arrays: dict


def make_problem(req: Request) -> Problem:
    """Synthetic measurements from the data set recipes of
    :mod:`ptychox_amd.synthetic` at the request's size, generated on the CPU,
    with the reference drivers' initial guesses: a uniform random real object
    and an aperture-only probe (python/train_ptychography.py:495-498), a flat
    object and aperture-only pupil for Fourier, ones for the coded object and
    surface (python/train_coded.py:172-177)."""
    from . import synthetic as syn

    n, V, seed = req.size, req.views, req.seed
    o = object_side(req)
    if req.kind == "conventional":
        side, step, origin, jitter = _conventional_geometry(n, o, V)
        d = syn.conventional_dataset(o, n, n / 4, side, step, origin, jitter, 1e5, 4.0, "cpu", seed=seed)
        init = np.random.default_rng(seed + 1).random((o, o), dtype=np.float32)
        return Problem("conventional", {
            "object": torch.from_numpy(init).to(torch.complex64),
            "aux": (d.probe.abs() > 0).to(torch.complex64),
            "positions": d.positions[:V].contiguous(),
            "measured": d.measured[:V].contiguous(),
        })
    if req.kind == "fourier":
        side = int(math.ceil(math.sqrt(V)))
        d = syn.fpm_dataset(o, n, 0.375 * n, side, max(n // 8, 1), "cpu", seed=seed)
        return Problem("fourier", {
            "object": torch.full((o, o), 0.6, dtype=torch.complex64),
            "aux": (d.pupil.abs() > 0).to(torch.complex64),
            "positions": d.leds[:V].contiguous(),
            "measured": d.measured[:V].contiguous(),
        })
    if req.kind == "nearfield":
        spread = max(min((o - n) // 2 - 1, 32), 0)
        d = syn.nearfield_dataset(o, n, V, spread, 20.0, 1e4, "cpu", seed=seed)
        return Problem("nearfield", {
            "object": torch.ones((o, o), dtype=torch.complex64),
            "aux": d.probe.clone(),
            "transfer": d.transfer,
            "positions": d.positions,
            "measured": d.measured,
        })
    side = int(math.ceil(math.sqrt(V)))
    d = syn.coded_dataset(n, side, 8, (n / 2) ** 2 / 20.0, 1e4, "cpu", seed=seed)
    return Problem("coded", {
        "object": torch.ones((n, n), dtype=torch.complex64),
        "aux": torch.ones((n, n), dtype=torch.complex64),
        "transfer": d.transfer_1,
        "shifts": d.shifts[:V].contiguous(),
        "measured": d.measured[:V].contiguous(),
    })


def pin_problem(prob: Problem) -> Problem:
    """The same request data in page-locked host memory, as a receiving
    server would hold it: the admission's uploads then run asynchronously on
    the request's stream instead of blocking the serving loop."""
    return Problem(prob.kind, {k: v.pin_memory() for k, v in prob.arrays.items()})


def _workspace_bytes(req: Request) -> int:
    s = shapes(req)
    (H, W), (h, w), V = s["object"], s["aux"], s["views"]
    if req.kind == "conventional":
        return _lib.query("ptx_base_step_workspace", H, W, h, w, V, 0)
    if req.kind == "fourier":
        return _lib.query("ptx_fpm_step_workspace", H, W, h, w, V, 0)
    if req.kind == "nearfield":
        return _lib.query("ptx_nearfield_step_workspace", H, W, h, w, V, 0)
    return _lib.query("ptx_coded_workspace", H, W, V, 1, 1)


_ALLOC_ROUND = 2 << 20  # the caching allocator's granularity for large blocks


def _rounded(nbytes: int) -> int:
    return _ALLOC_ROUND * ((int(nbytes) + _ALLOC_ROUND - 1) // _ALLOC_ROUND)


def footprint(req: Request) -> int:
    """HBM bytes a request holds while resident, from its shapes alone: each
    tensor it allocates, rounded up to the allocator's block size."""
    s = shapes(req)
    obj = s["object"][0] * s["object"][1]
    aux = s["aux"][0] * s["aux"][1]
    frame = s["frame"][0] * s["frame"][1]
    V = s["views"]
    tensors = [
        8 * obj, 8 * aux,              # unknowns
        8 * obj, 8 * aux,              # gradients
        4 * 2 * obj, 4 * 2 * obj,      # Adam moments of the object (per real component)
        4 * 2 * aux, 4 * 2 * aux,      # Adam moments of the auxiliary unknown
        4 * V * frame,                 # measured amplitudes
        _workspace_bytes(req),
        4 * V,                         # per-view losses
        4 * 4 * (req.epochs + 1),      # Adam step-size table
        8,                             # epoch counter
    ]
    if req.kind in ("conventional", "fourier", "nearfield"):
        tensors.append(8 * 2 * V)      # positions
    if req.kind in ("nearfield", "coded"):
        tensors.append(8 * frame)      # transfer function
    if req.kind == "coded":
        tensors.append(4 * 2 * V)      # shifts
    return sum(_rounded(t) for t in tensors)


def autotune_classes(requests: list[Request], problems: dict[int, Problem], device="cuda", reps: int = 10,
                     rounds: int = 2) -> dict:
    """Tune the FFT schedule (:mod:`ptychox_amd.autotune`) for the request
    classes in ``requests``: one representative per (kind, size) -- the one
    with the most views -- and, as the objective, the time of one fused step
    of each in turn, since the classes share table entries (a 512-point row
    pass serves near-field and coded requests alike).  The tables are
    process-wide: tune before serving, not while requests run (footprints and
    workspaces are sized from the tables in force).  Returns the autotuner's
    result (table, step-sum time before and after)."""
    from . import autotune as at

    dev = torch.device(device)
    rep: dict = {}
    for r in requests:
        key = (r.kind, r.size)
        if key not in rep or r.views > rep[key].views:
            rep[key] = r
    stream = torch.cuda.Stream(device=dev)
    jobs = [Job(r, problems[r.rid], dev, stream, graph=False) for r in rep.values()]
    keys: set = set()
    for r in rep.values():
        sh = shapes(r)
        if r.kind in ("conventional", "nearfield"):
            keys |= set(at.keys_for_scan(sh["aux"]))
        elif r.kind == "fourier":
            keys |= set(at.keys_for_fpm(sh["object"], sh["aux"]))
        else:
            keys |= set(at.keys_for_coded(sh["object"]))

    def run():
        # bracketed by the timing events the autotuner records on the current stream
        stream.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(stream):
            for j in jobs:
                j._fused_step()
        torch.cuda.current_stream(dev).wait_stream(stream)

    try:
        return at.autotune(run, sorted(keys), reps=reps, rounds=rounds)
    finally:
        torch.cuda.synchronize(dev)
        for j in jobs:
            j.release()


# ----------------------------------------------------------------------------- device-resident job

def hyper_table(lr: float, epochs: int) -> torch.Tensor:
    """Adam step sizes of a request, one row per epoch: row e holds
    ``adam_hyper`` of epoch e + 1 (cosine learning rate, one step per epoch),
    then a pad column.  A last row ``(0, 1, 1, 0)`` -- step size 0, with a
    bias correction and eps that keep the denominator >= 1 -- leaves the
    parameters unchanged if an epoch ever runs past the plan."""
    rows = [list(adam_hyper(cosine_lr(lr, 0.0, epochs, e), 0.9, 0.999, 1e-16, e)) + [0.0]
            for e in range(1, epochs + 1)]
    return torch.tensor(rows + [[0.0, 1.0, 1.0, 0.0]], dtype=torch.float32)


def epoch_hyper(table: torch.Tensor, counter: torch.Tensor, epochs: int,
                out: tuple[torch.Tensor, torch.Tensor] | None = None) -> torch.Tensor:
    """The row of ``table`` for the epoch ``counter`` (a one-element int64
    device tensor) counts, selected on the device; the index is clamped to
    the last row, so a counter past the plan never reads beyond the table.
    ``out`` = (index [1] int64, row [1, 4] float32) preallocated on the
    table's device: the selection then allocates nothing (a captured epoch
    that allocates nothing leaves its graph no memory pool to free)."""
    if out is None:
        return table.index_select(0, counter.clamp_max(epochs))[0]
    idx, row = out
    torch.clamp(counter, max=epochs, out=idx)
    torch.index_select(table, 0, idx, out=row)
    return row[0]


_HYPER_CACHE: dict = {}


def device_hyper_table(lr: float, epochs: int, device: torch.device) -> torch.Tensor:
    """``hyper_table(lr, epochs)`` on ``device``, built once per (lr, epochs,
    device) and shared read-only by the requests that use it.  Requests read
    it on their own streams, so the upload is complete before it is handed
    out (a synchronisation of the uploading stream, once per table)."""
    key = (float(lr), int(epochs), str(device))
    t = _HYPER_CACHE.get(key)
    if t is None:
        t = hyper_table(lr, epochs).to(device)
        torch.cuda.current_stream(device).synchronize()
        _HYPER_CACHE[key] = t
    return t


class Job:
    """One admitted request: its device state, stream and captured epoch."""

    def __init__(self, req: Request, prob: Problem, device: torch.device, stream: torch.cuda.Stream,
                 graph: bool = True, depth: int = 2):
        from .base import BaseStep
        from .coded import CodedStep
        from .fpm import FpmStep
        from .nearfield import NearFieldStep

        self.req, self.device, self.stream = req, device, stream
        self.graph_enabled, self.depth = bool(graph), int(depth)
        self.graph: torch.cuda.CUDAGraph | None = None
        self.launched = 0
        self.inflight: collections.deque = collections.deque()
        self.done_event: torch.cuda.Event | None = None
        before = torch.cuda.memory_allocated(device)
        with torch.cuda.stream(stream):
            # asynchronous on the request's stream when the host arrays are
            # pinned (pin_problem), so admission does not wait for the upload
            a = {k: v.to(device, non_blocking=True) for k, v in prob.arrays.items()}
            self.obj = a.pop("object")
            self.aux = a.pop("aux")
            self.a = a
            self.g_obj = torch.empty_like(self.obj)
            self.g_aux = torch.empty_like(self.aux)
            self.opt_o = AdamState(self.obj)
            # coded surface: real, non-negative after every step (python/train_coded.py:240-241)
            self.opt_a = AdamState(self.aux, real_nonneg=req.kind == "coded")
            V = a["measured"].shape[0]
            self.views = V
            if req.kind == "conventional":
                self.step = BaseStep(self.obj.shape, self.aux.shape, V, device, scatter="segments")
            elif req.kind == "fourier":
                self.step = FpmStep(self.obj.shape, self.aux.shape, V, device, scatter="segments")
            elif req.kind == "nearfield":
                self.step = NearFieldStep(self.obj.shape, self.aux.shape, V, device, scatter="segments")
            else:
                self.step = CodedStep(self.obj.shape, V, 1, device)
            self.hyper_table = device_hyper_table(req.lr, req.epochs, device)
            self.counter = torch.zeros(1, dtype=torch.int64, device=device)
            self._hyper_out = (torch.empty(1, dtype=torch.int64, device=device),
                               torch.empty((1, 4), dtype=torch.float32, device=device))
            self._loss_dev = torch.empty((), dtype=torch.float32, device=device)
        # the final loss, copied to pinned host memory behind the last epoch
        self._loss_host = torch.empty((), dtype=torch.float32, pin_memory=True)
        self.resident_bytes = torch.cuda.memory_allocated(device) - before

    def _fused_step(self) -> None:
        a, kind = self.a, self.req.kind
        if kind in ("conventional", "fourier"):
            self.step(self.obj, self.aux, a["positions"], a["measured"], self.g_obj, self.g_aux)
        elif kind == "nearfield":
            self.step(self.obj, self.aux, a["transfer"], a["positions"], a["measured"], self.g_obj, self.g_aux)
        else:
            self.step(self.obj, self.aux, a["transfer"], a["transfer"], a["shifts"], a["measured"], self.g_obj,
                      self.g_aux)

    def _launch_epoch(self) -> None:
        # The step sizes of the current epoch, selected on the device.  The
        # table has epochs + 1 rows (the last one all zeros); the index is
        # clamped on the device too, so even a replay past the planned epochs
        # reads the zero row instead of memory past the table.
        hyper = epoch_hyper(self.hyper_table, self.counter, self.req.epochs, self._hyper_out)
        self._fused_step()
        self.opt_o.step(self.g_obj, hyper)
        self.opt_a.step(self.g_aux, hyper)
        self.counter.add_(1)

    def _capture(self) -> None:
        g = torch.cuda.CUDAGraph()
        # capture on the request's stream without a device-wide synchronisation,
        # so the other requests keep running meanwhile
        g.capture_begin(capture_error_mode="thread_local")
        try:
            self._launch_epoch()
        finally:
            g.capture_end()
        self.graph = g

    def launch_next(self) -> None:
        """Queue the request's next epoch on its stream: eagerly for the
        first, then as a captured graph (captured at the second epoch and
        replayed after).  Raises RuntimeError, before any launch, once the
        request has run all its epochs."""
        if self.launched >= self.req.epochs:
            raise RuntimeError(f"request {self.req.rid} has already run its {self.req.epochs} epochs")
        with torch.cuda.stream(self.stream):
            if not self.graph_enabled or self.launched == 0:
                self._launch_epoch()
            else:
                if self.graph is None:
                    self._capture()
                self.graph.replay()
            if self.launched + 1 == self.req.epochs:
                # the final loss travels to the host behind the last epoch, so
                # reading it after completion needs no device synchronisation
                torch.mean(self.step.loss[: self.views], dim=0, out=self._loss_dev)
                self._loss_host.copy_(self._loss_dev, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.inflight.append(ev)
        self.launched += 1
        if self.launched == self.req.epochs:
            self.done_event = ev

    def pump(self) -> bool:
        """Queue epochs until ``depth`` are in flight; True if any was queued."""
        while self.inflight and self.inflight[0].query():
            self.inflight.popleft()
        queued = False
        while self.launched < self.req.epochs and len(self.inflight) < self.depth:
            self.launch_next()
            queued = True
        return queued

    @property
    def finished(self) -> bool:
        return self.done_event is not None and self.done_event.query()

    def final_loss(self) -> float:
        """Amplitude MSE at the last epoch's forward pass (the per-view losses
        are means over each frame's pixels); read once the request has
        finished (its copy to the host is done then)."""
        if not self.finished:
            raise RuntimeError(f"request {self.req.rid} has not finished")
        return float(self._loss_host)

    def take_graph(self) -> torch.cuda.CUDAGraph | None:
        """Detach the captured epoch (to be destroyed later by the caller)."""
        g, self.graph = self.graph, None
        return g

    def release(self) -> None:
        """Drop the graph and every device buffer now (all queued work is done)."""
        self.graph = None
        self.step = self.opt_o = self.opt_a = None
        self.obj = self.aux = self.g_obj = self.g_aux = None
        self.a = {}
        self.hyper_table = self.counter = self._hyper_out = self._loss_dev = None
        self.inflight.clear()
        self.done_event = None


@dataclass
class Completed:
    rid: int
    kind: str
    arrival: float
    admitted: float
    finished: float
    final_loss: float
    footprint: int = 0
    resident: int = 0
    object: torch.Tensor | None = None
    aux: torch.Tensor | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class Server:
    """``run(requests)`` replays the arrival schedule in real time on one GPU
    and returns one :class:`Completed` record per request, in request order."""

    def __init__(self, device="cuda", hbm_budget: float | None = None, slots: int = 8, graph: bool = True,
                 depth: int = 2, keep_results: bool = False, retire_batch: int = 32):
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            hbm_budget = 0.8 * free
        self.budget = int(hbm_budget)
        self.slots, self.graph, self.depth = int(slots), bool(graph), int(depth)
        self.keep = bool(keep_results)
        # Destroying a graph waits for the work in flight on the device, so
        # finished requests' graphs are destroyed together: when this many
        # have piled up, or whenever no request is running.  Their captures
        # allocate nothing, so a retired graph holds no device buffers.
        self.retire_batch = max(int(retire_batch), 1)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.slots)]
        self.peak_in_flight = 0
        self.peak_bytes = 0

    def run(self, requests: list[Request], problems: dict[int, Problem] | None = None) -> list[Completed]:
        _lib.lib()
        if problems is None:
            problems = {r.rid: make_problem(r) for r in requests}
        for r in requests:
            if footprint(r) > self.budget:
                raise ValueError(f"request {r.rid} needs {footprint(r)} bytes, more than the HBM budget {self.budget}")
        pending = collections.deque(sorted(requests))
        queue: list[Request] = []
        active: dict[int, tuple[Job, int, float]] = {}  # slot -> (job, footprint, admission time)
        free_slots = list(range(self.slots))
        used = 0
        done: list[Completed] = []
        retired: list = []  # graphs of finished requests, destroyed in batches
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        while pending or queue or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            # retire finished requests
            for slot in list(active):
                job, nbytes, admitted = active[slot]
                if job.finished:
                    fin = time.perf_counter() - t0
                    # read the results on the request's own stream: no ordering
                    # against the other requests' work
                    with torch.cuda.stream(job.stream):
                        rec = Completed(job.req.rid, job.req.kind, job.req.arrival, admitted, fin,
                                        job.final_loss(), nbytes, job.resident_bytes)
                        if self.keep:
                            rec.object, rec.aux = job.obj.clone(), job.aux.clone()
                    retired.append(job.take_graph())
                    job.release()
                    done.append(rec)
                    del active[slot]
                    free_slots.append(slot)
                    used -= nbytes
            # admit in arrival order while a slot is free and the footprint fits
            while queue and free_slots:
                nbytes = footprint(queue[0])
                if used + nbytes > self.budget:
                    break
                req = queue.pop(0)
                slot = free_slots.pop(0)
                job = Job(req, problems[req.rid], self.device, self.streams[slot], self.graph, self.depth)
                active[slot] = (job, nbytes, time.perf_counter() - t0)
                used += nbytes
            if retired and (len(retired) >= self.retire_batch or not active):
                retired.clear()
            self.peak_in_flight = max(self.peak_in_flight, len(active))
            self.peak_bytes = max(self.peak_bytes, used)
            progressed = False
            for job, _, _ in active.values():
                progressed |= job.pump()
            if not progressed:
                if active:
                    time.sleep(2e-5)
                elif pending and not queue:
                    time.sleep(max(0.0, min(1e-3, pending[0].arrival - (time.perf_counter() - t0))))
        return sorted(done, key=lambda c: c.rid)


def run_plain(requests: list[Request], problems: dict[int, Problem], device="cuda") -> list[Completed]:
    """The plain per-request path: one request at a time in arrival order, the
    reference's unfused torch operators, ``F.mse_loss``, ``loss.backward()``
    and ``torch.optim.Adam`` -- the reference driver's loop with one full-data
    step per epoch."""
    from . import ptychoX_base, ptychoX_coded, ptychoX_fourier, ptychoX_nearfield

    dev = torch.device(device)
    done = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for req in sorted(requests):
        now = time.perf_counter() - t0
        if req.arrival > now:
            time.sleep(req.arrival - now)
        start = time.perf_counter() - t0
        a = {k: v.to(dev) for k, v in problems[req.rid].arrays.items()}
        obj = torch.nn.Parameter(a["object"])
        aux = torch.nn.Parameter(a["aux"])
        opt = torch.optim.Adam([obj, aux], lr=req.lr, betas=(0.9, 0.999), eps=1e-16)

        def forward():
            if req.kind == "conventional":
                return ptychoX_base(obj, aux, a["positions"])
            if req.kind == "fourier":
                return ptychoX_fourier(obj, aux, a["positions"])
            if req.kind == "nearfield":
                return ptychoX_nearfield(obj, aux, a["transfer"], a["positions"])
            return ptychoX_coded(obj, aux, a["transfer"], a["transfer"], a["shifts"], 1)

        loss = None
        for e in range(1, req.epochs + 1):
            for g in opt.param_groups:
                g["lr"] = cosine_lr(req.lr, 0.0, req.epochs, e)
            loss = torch.nn.functional.mse_loss(forward(), a["measured"])
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            if req.kind == "coded":
                with torch.no_grad():
                    aux.copy_(aux.real.clamp_min(0.0).to(torch.complex64))
        fl = float(loss.detach())
        done.append(Completed(req.rid, req.kind, req.arrival, start, time.perf_counter() - t0, fl))
    return sorted(done, key=lambda c: c.rid)


def summarize(done: list[Completed]) -> dict:
    """p50 / p99 / mean latency (arrival to completion) and throughput over
    the span from the first arrival to the last completion."""
    lat = np.array([c.latency for c in done]) * 1e3
    span = max(c.finished for c in done) - min(c.arrival for c in done)
    return {
        "requests": len(done),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3),
        "span_s": round(span, 3),
        "throughput_req_per_s": round(len(done) / span, 3),
    }

"""The conventional training step with the Adam updates fused into the
gradient epilogues (``ptx_base_train_step``) against the unfused sequence the
reference driver runs per mini-batch (python/train_ptychography.py:549-559:
step -> ``optimizer.step()``), i.e. ``ptx_base_step_indexed`` followed by
``ptx_adam_step`` on the sample and on the probe.  Both compute the same
gradient sums in the same order and the same Adam arithmetic, so sample,
probe, both moments and the per-view losses must agree bitwise, over two
consecutive steps (the second with non-zero moments), on each gather shape:
the single-pass gather, the segmented gather with tiles of several segments
(combine pass), and the whole-frame path (32^2).  The atomic scatter mode
adds the object gradient in whatever order the atomics land, so two runs of
either sequence differ in the last bits: there the probe (a fixed-order sum)
must still agree bitwise and the sample and its moments to 1e-6 norm-wise.
The reconstruction loop is compared with ``fused_adam`` on and off, graph
replays included."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from tests.helpers import crandn, disc_probe, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

CASES = [  # H, h, views in the table, batch
    (256, 64, 30, 12),   # single-pass gather
    (160, 64, 90, 64),   # segmented gather, tiles of several segments
    (80, 32, 40, 20),    # whole-frame step
]


def _problem(gpu, H, h, count, seed):
    rng = np.random.default_rng(seed)
    s = to_dev((smooth_object(rng, H) + 0.05 * crandn(rng, H, H)).astype(np.complex64), gpu)
    p = to_dev((disc_probe(h, h / 4) + 0.02 * crandn(rng, h, h)).astype(np.complex64), gpu)
    pos = to_dev(rng.integers(0, H - h + 1, (count, 2)).astype(np.int64), gpu)
    y = to_dev(rng.uniform(0.0, 2.0, (count, h, h)).astype(np.float32), gpu)
    return s, p, pos, y


class _State:
    def __init__(self, s, p):
        self.s, self.p = s.clone(), p.clone()
        self.m = [torch.zeros_like(torch.view_as_real(t)) for t in (s, p)]
        self.v = [torch.zeros_like(torch.view_as_real(t)) for t in (s, p)]


def _hyper(gpu, t, lr):
    from ptychox_amd.train import adam_hyper

    return torch.tensor(adam_hyper(lr, 0.9, 0.999, 1e-16, t), dtype=torch.float32, device=gpu)


def _unfused(lib, st, pos, y, order, B, hyp, ws, flags, gpu):
    H, h = st.s.shape[0], st.p.shape[0]
    gs, gp = torch.empty_like(st.s), torch.empty_like(st.p)
    loss = torch.empty(B, dtype=torch.float32, device=gpu)
    lib.call("ptx_base_step_indexed", lib.ptr(st.s), H, H, lib.ptr(st.p), h, h, lib.ptr(pos), lib.ptr(y),
             pos.shape[0], lib.ptr(order), B, 1.0 / (B * h * h), lib.ptr(loss), None, lib.ptr(gs), lib.ptr(gp),
             lib.ptr(ws), ws.numel(), flags, lib.stream_of(gpu))
    for k, (x, g) in enumerate(((st.s, gs), (st.p, gp))):
        lib.call("ptx_adam_step", lib.ptr(x), lib.ptr(g), lib.ptr(st.m[k]), lib.ptr(st.v[k]), st.m[k].numel(),
                 0.9, 0.999, lib.ptr(hyp[k]), 0, lib.stream_of(gpu))
    return loss, gp


def _fused(lib, st, pos, y, order, B, hyp, ws, flags, gpu, update_probe=True):
    H, h = st.s.shape[0], st.p.shape[0]
    gs, gp = torch.empty_like(st.s), torch.empty_like(st.p)
    loss = torch.empty(B, dtype=torch.float32, device=gpu)
    ao = lib.PtxAdam(st.m[0].data_ptr(), st.v[0].data_ptr(), hyp[0].data_ptr(), 0.9, 0.999)
    ap = lib.PtxAdam(st.m[1].data_ptr(), st.v[1].data_ptr(), hyp[1].data_ptr(), 0.9, 0.999)
    lib.call("ptx_base_train_step", lib.ptr(st.s), H, H, lib.ptr(st.p), h, h, lib.ptr(pos), lib.ptr(y),
             pos.shape[0], lib.ptr(order), B, 1.0 / (B * h * h), lib.ptr(loss), None, lib.ptr(gs), lib.ptr(gp),
             ctypes.c_void_p(ctypes.addressof(ao)),
             ctypes.c_void_p(ctypes.addressof(ap)) if update_probe else None,
             lib.ptr(ws), ws.numel(), flags, lib.stream_of(gpu))
    return loss, gp


@pytest.mark.parametrize("H,h,count,B", CASES)
@pytest.mark.parametrize("scatter", ["segments", "atomic"])
def test_fused_adam_step_equals_step_then_adam(gpu, H, h, count, B, scatter):
    from ptychox_amd import _lib as lib
    from ptychox_amd._utils import scatter_flags

    flags = scatter_flags(scatter)
    s, p, pos, y = _problem(gpu, H, h, count, seed=H + B)
    ws = lib.workspace(lib.query("ptx_base_step_workspace", H, H, h, h, B, flags), gpu)
    a, b = _State(s, p), _State(s, p)
    g = torch.Generator(device="cpu").manual_seed(B)
    for t in (1, 2):
        order = torch.randperm(count, generator=g)[:B].to(gpu)
        hyp = [_hyper(gpu, t, 1e-2), _hyper(gpu, t, 5e-3)]
        la, _ = _unfused(lib, a, pos, y, order, B, hyp, ws, flags, gpu)
        lb, _ = _fused(lib, b, pos, y, order, B, hyp, ws, flags, gpu)
        torch.cuda.synchronize()
        print(f"fused Adam {H}/{h} B{B} {scatter} step {t}: max |ds| "
              f"{float((a.s - b.s).abs().max()):.3e} |dp| {float((a.p - b.p).abs().max()):.3e}", flush=True)
        assert not torch.equal(a.s, s)  # the sample moved
        if scatter == "atomic":
            # step 1 starts from equal unknowns: equal losses and probe update
            assert t > 1 or (torch.equal(la, lb) and torch.equal(a.p, b.p))
            assert nrel(la, lb) < 1e-6, t
            assert nrel(a.s, b.s) < 1e-6 and nrel(a.p, b.p) < 1e-6, t
            for k in range(2):
                assert nrel(a.m[k], b.m[k]) < 1e-6 and nrel(a.v[k], b.v[k]) < 1e-6, (t, k)
            continue
        assert torch.equal(la, lb), t
        assert torch.equal(a.s, b.s) and torch.equal(a.p, b.p), t
        for k in range(2):
            assert torch.equal(a.m[k], b.m[k]) and torch.equal(a.v[k], b.v[k]), (t, k)


def test_fused_adam_step_without_probe_update_stores_its_gradient(gpu):
    from ptychox_amd import _lib as lib

    H, h, count, B = 160, 64, 90, 64
    s, p, pos, y = _problem(gpu, H, h, count, seed=3)
    ws = lib.workspace(lib.query("ptx_base_step_workspace", H, H, h, h, B, 0), gpu)
    a, b = _State(s, p), _State(s, p)
    order = torch.randperm(count, generator=torch.Generator().manual_seed(0))[:B].to(gpu)
    hyp = [_hyper(gpu, 1, 1e-2), _hyper(gpu, 1, 5e-3)]
    _, gp_ref = _unfused(lib, a, pos, y, order, B, hyp, ws, 0, gpu)
    _, gp = _fused(lib, b, pos, y, order, B, hyp, ws, 0, gpu, update_probe=False)
    torch.cuda.synchronize()
    assert torch.equal(a.s, b.s)
    assert torch.equal(b.p, p)  # not updated
    assert torch.equal(gp, gp_ref)  # its gradient stored instead


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("scatter", ["segments", "atomic"])
def test_reconstruction_fused_equals_unfused(gpu, graph, scatter):
    from ptychox_amd.train import ConventionalReconstruction

    H, h, count, batch = 192, 64, 50, 16  # four mini-batches, the last one partial
    s, p, pos, y = _problem(gpu, H, h, count, seed=11)
    runs = {}
    for fused in (False, True):
        gen = torch.Generator(device=gpu).manual_seed(5)
        rec = ConventionalReconstruction(s, p, pos, y, batch, lr=1e-2, probe_lr=5e-3, epochs=3, generator=gen,
                                         graph=graph, scatter=scatter, fused_adam=fused)
        losses = [r.loss for r in rec.fit()]
        torch.cuda.synchronize()
        runs[fused] = (losses, rec.sample.clone(), rec.probe.clone())
    print(f"reconstruction {scatter} graph={graph}: epoch losses unfused {runs[False][0]} fused {runs[True][0]}",
          flush=True)
    if scatter == "atomic":  # the atomics' order differs from run to run
        assert max(abs(a - b) / a for a, b in zip(runs[False][0], runs[True][0])) < 1e-5
        assert nrel(runs[False][1], runs[True][1]) < 1e-5 and nrel(runs[False][2], runs[True][2]) < 1e-5
        return
    assert runs[False][0] == runs[True][0]
    assert torch.equal(runs[False][1], runs[True][1])
    assert torch.equal(runs[False][2], runs[True][2])

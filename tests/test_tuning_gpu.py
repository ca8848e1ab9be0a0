"""Every operator stays in parity with the FP64 oracle under both compiled
FFT schedules (8 or 16 elements per thread in every pass kind) and under
extreme work-group targets of the view-walking passes, and the autotuner
leaves a valid table behind."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import crandn, disc_probe, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5
SIZES = (16, 32, 64, 128, 256, 512, 1024, 2048)


def _force(elems, groups=0):
    from ptychox_amd import _lib

    for kind in range(5):
        for n in SIZES:
            _lib.call("ptx_tune_set", kind, n, elems)
            _lib.call("ptx_tune_set_groups", kind, n, groups)


# (elements per thread, work-group target): target 1 makes every work-group of
# the view-walking passes walk all views of its rows / columns (the pipelined
# view loops, with the alternating LDS regions, at these small shapes); 65536
# leaves one view per work-group
@pytest.fixture(params=[(8, 0), (16, 0), (8, 1), (16, 1), (8, 65536)])
def schedule(request, gpu):
    _force(*request.param)
    yield request.param
    _force(0, 0)


def test_all_operators_under_schedule(gpu, schedule):
    from ptychox_amd import ptychoX_base, ptychoX_coded, ptychoX_nearfield

    rng = np.random.default_rng(schedule[0] + schedule[1])
    # conventional
    s = smooth_object(rng, 160)
    p = disc_probe(64, 16)
    pos = rng.integers(0, 97, (5, 2)).astype(np.int64)
    g = rng.standard_normal((5, 64, 64)).astype(np.float32)
    st, pt = to_dev(s, gpu).requires_grad_(True), to_dev(p, gpu).requires_grad_(True)
    ptychoX_base(st, pt, to_dev(pos, gpu), use_cuda_kernels=True).backward(to_dev(g, gpu))
    f = po.base_forward(s, p, pos)
    gs, gp = po.base_backward(g, s.shape, p, f["patches"], f["far_field"], pos)
    assert max(nrel(st.grad, gs), nrel(pt.grad, gp)) < TOL
    # near-field
    t = np.exp(1j * rng.uniform(0, 6.28, (64, 64))).astype(np.complex64)
    st, pt = to_dev(s, gpu).requires_grad_(True), to_dev(p, gpu).requires_grad_(True)
    ptychoX_nearfield(st, pt, to_dev(t, gpu), to_dev(pos, gpu), use_cuda_kernels=True).backward(to_dev(g, gpu))
    f = po.nearfield_forward(s, p, t, pos)
    gs, gp = po.nearfield_backward(g, s.shape, p, t, f["patches"], f["field"], pos)
    assert max(nrel(st.grad, gs), nrel(pt.grad, gp)) < TOL
    # FPM
    o = (smooth_object(rng, 128) + 0.05 * crandn(rng, 128, 128)).astype(np.complex64)
    q = crandn(rng, 32, 32)
    led = np.array([[40, 40], [48, 44], [-3, 120]], dtype=np.int64)
    gf = rng.standard_normal((3, 32, 32)).astype(np.float32)
    # forward against the oracle; the backward against the oracle's backward
    # applied to the HIP forward's own image field: a random upstream gradient
    # weights the phase E/|E| of near-zero pixels, where the FP32 field's
    # rounding (any transform's) would otherwise dominate the comparison
    from ptychox_amd.fpm import native_backward, native_forward

    f = po.fpm_forward(o, q, led)
    od, qd, ld = to_dev(o, gpu), to_dev(q, gpu), to_dev(led, gpu)
    spec, field, amp = native_forward(od, qd, ld)
    go_n, gq_n = native_backward(to_dev(gf, gpu), field, spec, qd, ld)
    field64 = field.cpu().numpy().astype(np.complex128)
    go, gq = po.fpm_backward(gf, o.shape, q, f["sub_spectrum"], field64, led)
    e = (nrel(amp, f["amplitude"]), nrel(field, f["image_field"]), nrel(go_n, go), nrel(gq_n, gq))
    print(f"schedule (E, groups)={schedule}: fourier amp {e[0]:.2e} field {e[1]:.2e} grad_object {e[2]:.2e} "
          f"grad_pupil {e[3]:.2e}", flush=True)
    assert max(e) < TOL
    # coded
    c = crandn(rng, 64, 64)
    t1, t2 = crandn(rng, 64, 64), crandn(rng, 64, 64)
    sh = np.array([[0.01, -0.02], [0.03, 0.01], [-0.05, 0.02]], dtype=np.float32)
    o = (smooth_object(rng, 64) + 0.05 * crandn(rng, 64, 64)).astype(np.complex64)
    gc = rng.standard_normal((3, 32, 32)).astype(np.float32)
    ot, ct = to_dev(o, gpu).requires_grad_(True), to_dev(c, gpu).requires_grad_(True)
    ptychoX_coded(ot, ct, to_dev(t1, gpu), to_dev(t2, gpu), to_dev(sh, gpu), 2,
                  use_cuda_kernels=True).backward(to_dev(gc, gpu))
    f = po.coded_forward(o, c, t1, t2, sh.astype(np.float64), 2)
    go, gcc = po.coded_backward(gc, f, c, t1, t2, 2)
    e = (nrel(ot.grad, go), nrel(ct.grad, gcc))
    print(f"schedule (E, groups)={schedule}: coded {e[0]:.2e} {e[1]:.2e}", flush=True)
    assert max(e) < TOL


def test_autotune_picks_valid_table(gpu):
    from ptychox_amd import autotune
    from ptychox_amd.base import BaseStep

    rng = np.random.default_rng(0)
    s = to_dev(smooth_object(rng, 512), gpu)
    p = to_dev(disc_probe(128, 32), gpu)
    pos = to_dev(rng.integers(0, 385, (16, 2)).astype(np.int64), gpu)
    y = torch.rand(16, 128, 128, device=gpu)
    gs, gp = torch.empty_like(s), torch.empty_like(p)
    step = BaseStep(s.shape, p.shape, 16, gpu)
    try:
        keys = autotune.keys_for_scan((128, 128))
        res = autotune.autotune(lambda: step(s, p, pos, y, gs, gp), keys, reps=3)
        t = res["table"]
        assert {v for k, v in t.items() if not k.endswith(autotune.GROUPS_SUFFIX)} <= {8, 16}
        for kind, n in keys:
            assert t[f"{autotune.KIND_NAMES[kind]}:{n}{autotune.GROUPS_SUFFIX}"] in autotune.GROUP_CANDIDATES[kind]
        ntrials = sum(len(autotune.CANDIDATES) + len(autotune.GROUP_CANDIDATES[k]) for k, _ in keys)
        assert res["after_ms"] > 0 and len(res["trials"]) == ntrials
        # the tuned step still matches the default schedule's result
        l_tuned = step(s, p, pos, y, gs, gp).clone()
        g_tuned = gs.clone()
        _force(0, 0)
        l_def = step(s, p, pos, y, gs, gp).clone()
        assert nrel(l_tuned, l_def) < TOL and nrel(g_tuned, gs) < TOL
    finally:
        _force(0, 0)

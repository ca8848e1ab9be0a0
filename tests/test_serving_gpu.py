"""Serving loop on the GPU: each request's result is bitwise the same alone and
in a mix, admission keeps the resident footprint within the HBM budget, graph
replay equals eager launches, and the fused path tracks the plain torch loop."""

import pytest
import torch

from ptychox_amd import serving as sv

pytestmark = pytest.mark.gpu

MIX = [
    sv.Request(0.0, 0, "conventional", 64, 100, seed=11, epochs=6),
    sv.Request(0.0, 1, "conventional", 128, 49, seed=12, epochs=5),
    sv.Request(0.001, 2, "fourier", 64, 81, seed=13, epochs=6),
    sv.Request(0.002, 3, "nearfield", 128, 16, seed=14, epochs=4),
    sv.Request(0.002, 4, "coded", 128, 9, seed=15, epochs=5),
    sv.Request(0.003, 5, "conventional", 64, 30, seed=16, epochs=7),
    sv.Request(0.004, 6, "fourier", 32, 25, seed=17, epochs=3),
]


@pytest.fixture(scope="module")
def problems():
    return {r.rid: sv.make_problem(r) for r in MIX}


def _alone(gpu, req, prob, **kw):
    srv = sv.Server(gpu, hbm_budget=4e9, keep_results=True, **kw)
    return srv.run([req], {req.rid: prob})[0]


def test_alone_equals_mix(gpu, problems):
    srv = sv.Server(gpu, hbm_budget=4e9, slots=4, keep_results=True)
    mixed = srv.run(MIX, problems)
    assert [c.rid for c in mixed] == [r.rid for r in MIX]
    assert srv.peak_in_flight > 1
    for req, m in zip(MIX, mixed):
        a = _alone(gpu, req, problems[req.rid])
        print(req.kind, req.size, req.views, m.final_loss, a.final_loss, flush=True)
        assert torch.equal(m.object, a.object), req
        assert torch.equal(m.aux, a.aux), req
        assert m.final_loss == a.final_loss


def test_admission_respects_budget(gpu, problems):
    sizes = [sv.footprint(r) for r in MIX]
    budget = max(sizes) + min(sizes)
    srv = sv.Server(gpu, hbm_budget=budget, slots=8, keep_results=False)
    done = srv.run(MIX, problems)
    assert len(done) == len(MIX)
    assert srv.peak_bytes <= budget
    for c in done:
        assert 0 < c.resident <= c.footprint, c
        assert c.admitted >= c.arrival


def test_graph_equals_eager(gpu, problems):
    for req in (MIX[0], MIX[2], MIX[3], MIX[4]):
        g = _alone(gpu, req, problems[req.rid], graph=True)
        e = _alone(gpu, req, problems[req.rid], graph=False)
        assert torch.equal(g.object, e.object), req
        assert torch.equal(g.aux, e.aux), req


@pytest.mark.parametrize("idx", [0, 2, 3, 4])
def test_tracks_plain_torch_loop(gpu, problems, idx):
    # Adam with eps = 1e-16 normalises every gradient component to a step of
    # about lr, including components at rounding level, so parameter paths of
    # two correct implementations drift apart by O(lr) in a few pixels; the
    # losses along the path agree far more tightly.
    req = MIX[idx]
    fused = _alone(gpu, req, problems[req.rid])
    plain = sv.run_plain([req], problems, gpu)[0]
    rel = abs(fused.final_loss - plain.final_loss) / abs(plain.final_loss)
    print(req.kind, fused.final_loss, plain.final_loss, rel, flush=True)
    assert rel < 2e-3


def test_finished_job_refuses_another_epoch(gpu, problems):
    # the host guard: a finished request raises before anything is launched
    req = MIX[6]
    job = sv.Job(req, problems[req.rid], gpu, torch.cuda.Stream(device=gpu), graph=True, depth=req.epochs + 1)
    while job.launched < req.epochs:
        job.launch_next()
    torch.cuda.synchronize()
    obj, aux = job.obj.clone(), job.aux.clone()
    assert int(job.counter) == req.epochs
    with pytest.raises(RuntimeError, match="already run"):
        job.launch_next()
    torch.cuda.synchronize()
    assert int(job.counter) == req.epochs and job.launched == req.epochs
    assert torch.equal(job.obj, obj) and torch.equal(job.aux, aux)
    # the device clamp: a replay past the plan (bypassing the guard) reads the
    # table's last row, whose step size 0 leaves the unknowns unchanged
    with torch.cuda.stream(job.stream):
        job.graph.replay()
    torch.cuda.synchronize()
    assert int(job.counter) == req.epochs + 1
    assert torch.equal(job.obj, obj) and torch.equal(job.aux, aux)


def _reset_tuning():
    from ptychox_amd import _lib

    for kind in range(6):
        for n in (16, 32, 64, 128, 256, 512, 1024, 2048):
            _lib.call("ptx_tune_set", kind, n, 0)
            _lib.call("ptx_tune_set_groups", kind, n, 0)


def test_schedule_tuned_for_the_mix_keeps_alone_equal_mix(gpu, problems):
    # the mix's own schedule (one request per class, their steps' sum): every
    # setting it leaves is a compiled variant, and under it each request is
    # still bitwise the same alone and in the mix (footprints follow the table)
    try:
        res = sv.autotune_classes(MIX, problems, gpu, reps=3, rounds=1)
        print("tuned for the mix:", res["table"], res["before_ms"], res["after_ms"], flush=True)
        for key, value in res["table"].items():
            assert value in ((0, 8, 16) if "@" not in key else range(0, 65537)), (key, value)
        srv = sv.Server(gpu, hbm_budget=4e9, slots=4, keep_results=True)
        mixed = srv.run(MIX, problems)
        for req, m in zip(MIX, mixed):
            a = _alone(gpu, req, problems[req.rid])
            assert torch.equal(m.object, a.object), req
            assert torch.equal(m.aux, a.aux), req
            assert m.final_loss == a.final_loss
            assert 0 < m.resident <= m.footprint, m
    finally:
        _reset_tuning()

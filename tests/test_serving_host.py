"""Host side of the serving loop (no GPU): the request stream, the shapes and
HBM footprints used for admission, and the synthetic problems."""

import math

import numpy as np
import pytest
import torch

from ptychox_amd import serving as sv


def test_standard_mix_classes_and_arrivals():
    reqs = sv.standard_mix(400, 50.0, seed=2)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) >= 0)
    assert abs(len(reqs) / t[-1] - 50.0) < 10.0  # Poisson rate
    seen = {(r.kind, r.size) for r in reqs}
    assert seen == {(k, n) for k, n, _, _ in sv.REQUEST_CLASSES}
    for r in reqs:
        assert r.seed == r.rid and r.epochs == 20
        if r.kind == "conventional":
            assert 100 <= r.views <= 400
        elif r.kind == "fourier":
            assert r.views in (81, 225)
        elif r.kind == "nearfield":
            assert r.views == 64
        else:
            assert r.views == 25
    burst = sv.standard_mix(10, math.inf)
    assert all(r.arrival == 0.0 for r in burst)


def test_object_sides_follow_the_class_table():
    for kind, n, o, (lo, hi) in sv.REQUEST_CLASSES:
        s = sv.shapes(sv.Request(0.0, 0, kind, n, hi))
        assert s["object"] == (o, o) and s["aux"] == (n, n)
    with pytest.raises(ValueError):
        sv.shapes(sv.Request(0.0, 0, "holography", 64, 4))


@pytest.mark.parametrize("n,obj,views", [(128, 512, 100), (128, 512, 400), (256, 1024, 400), (64, 256, 1)])
def test_conventional_raster_fits(n, obj, views):
    side, step, origin, jitter = sv._conventional_geometry(n, obj, views)
    assert side * side >= views
    assert origin - jitter >= 0
    assert origin + step * (side - 1) + jitter + n <= obj


def test_footprint_bounds_the_buffers_and_grows_with_views():
    small = sv.Request(0.0, 0, "conventional", 128, 100)
    large = sv.Request(0.0, 1, "conventional", 128, 400)
    assert sv.footprint(large) > sv.footprint(small)
    for kind, n, o, (lo, hi) in sv.REQUEST_CLASSES:
        r = sv.Request(0.0, 0, kind, n, hi)
        fp = sv.footprint(r)
        # at least the measurements, the unknowns with gradients and moments, and the workspace
        floor = 4 * hi * n * n + 32 * (o * o + n * n) + sv._workspace_bytes(r)
        assert fp >= floor
        assert fp % (2 << 20) == 0


@pytest.mark.parametrize("kind,n,views", [("conventional", 64, 30), ("fourier", 32, 25), ("nearfield", 64, 9),
                                          ("coded", 64, 9)])
def test_problems_match_declared_shapes(kind, n, views):
    r = sv.Request(0.0, 3, kind, n, views, seed=3)
    p = sv.make_problem(r)
    s = sv.shapes(r)
    assert tuple(p.arrays["object"].shape) == s["object"]
    assert tuple(p.arrays["aux"].shape) == s["aux"]
    assert tuple(p.arrays["measured"].shape) == (views,) + s["frame"]
    assert p.arrays["object"].dtype == torch.complex64 and p.arrays["measured"].dtype == torch.float32
    if kind in ("conventional", "nearfield"):
        pos = p.arrays["positions"]
        assert pos.min() >= 0 and pos.max() + n <= s["object"][0]
    # deterministic from the request's seed
    q = sv.make_problem(r)
    assert torch.equal(p.arrays["measured"], q.arrays["measured"])


def test_step_size_table_index_is_clamped():
    # The table has one row per epoch plus a harmless last row; a counter past
    # the plan (a stray graph replay) reads that row, never beyond the table.
    from ptychox_amd.train import adam_hyper, cosine_lr

    epochs = 5
    table = sv.hyper_table(1e-2, epochs)
    assert table.shape == (epochs + 1, 4)
    for e in range(epochs):
        row = sv.epoch_hyper(table, torch.tensor([e]), epochs)
        want = adam_hyper(cosine_lr(1e-2, 0.0, epochs, e + 1), 0.9, 0.999, 1e-16, e + 1)
        assert torch.allclose(row[:3], torch.tensor(want, dtype=torch.float32))
    for c in (epochs, epochs + 1, epochs + 7, 10**6):
        assert torch.equal(sv.epoch_hyper(table, torch.tensor([c]), epochs), table[epochs])
    # Adam's arithmetic with the last row: step size 0 and a denominator >= 1,
    # so the parameter is unchanged whatever the moments hold (v = 0 included)
    step, bc2, eps = (float(x) for x in table[epochs, :3])
    p = torch.tensor([1.5, -2.0, 0.0])
    m = torch.tensor([0.3, 0.0, -1e30])
    v = torch.tensor([0.0, 4.0, 1e30])
    upd = p + (-step) * (m / (v.sqrt() / bc2 + eps))
    assert torch.equal(upd, p)

"""The serving mix at its configured shapes: one request of each class of
``serving.REQUEST_CLASSES`` (conventional probe 128 / object 512 and probe
256 / object 1024, Fourier 128^2 pupil / 512^2 object, near-field 512 /
1024, coded 512^2), 20 Adam epochs each.

* each request's result is bitwise the same alone and in the mix;
* the fused first epoch (loss per view, both gradients) agrees with the plain
  per-request path (the reference's torch operators, ``F.mse_loss``,
  autograd; python/train_ptychography.py:549-559) to <= 1e-5;
* final losses: the plain path whose first Adam step is fed the fused
  step's gradients ends within 1e-4 of the fused path's final loss, for every
  class.  Adam with eps = 1e-16 moves every component with |g| >> eps by a
  full learning rate at its first step, so components whose exact gradient is
  zero take a +-lr step from rounding-level noise (~1e-8 of the largest
  component) that one FP32 implementation returns and another rounds to an
  exact 0 -- the Fourier class's aperture pupil has such components.  This
  seeds a final-loss difference of ~1 % there; the seeded control shows the
  rest of the trajectory is the plain path's.  The unseeded gap is printed
  for every class and bounded by 1e-4 where no such components occur.
"""

from __future__ import annotations

import pytest
import torch

from ptychox_amd import serving as sv
from tests.helpers import nrel

pytestmark = pytest.mark.gpu

# one request of each class at its configured shapes (views within the class range)
MIX = [
    sv.Request(0.0, 1, "conventional", 128, 250, seed=1),
    sv.Request(0.0, 2, "conventional", 256, 400, seed=2),
    sv.Request(0.001, 3, "fourier", 128, 225, seed=3),
    sv.Request(0.002, 4, "nearfield", 512, 64, seed=4),
    sv.Request(0.002, 5, "coded", 512, 25, seed=5),
]
CLASSES_WITHOUT_EXACT_ZEROS = ("conventional", "nearfield", "coded")


@pytest.fixture(scope="module")
def problems():
    return {r.rid: sv.make_problem(r) for r in MIX}


def _forward(req, obj, aux, a):
    from ptychox_amd import ptychoX_base, ptychoX_coded, ptychoX_fourier, ptychoX_nearfield

    if req.kind == "conventional":
        return ptychoX_base(obj, aux, a["positions"])
    if req.kind == "fourier":
        return ptychoX_fourier(obj, aux, a["positions"])
    if req.kind == "nearfield":
        return ptychoX_nearfield(obj, aux, a["transfer"], a["positions"])
    return ptychoX_coded(obj, aux, a["transfer"], a["transfer"], a["shifts"], 1)


def _plain(req, prob, gpu, first_grads=None):
    """sv.run_plain's loop for one request; first_grads replaces the
    gradients of the first Adam step."""
    from ptychox_amd.train import cosine_lr

    a = {k: v.to(gpu) for k, v in prob.arrays.items()}
    obj = torch.nn.Parameter(a["object"].clone())
    aux = torch.nn.Parameter(a["aux"].clone())
    opt = torch.optim.Adam([obj, aux], lr=req.lr, betas=(0.9, 0.999), eps=1e-16)
    loss = None
    for e in range(1, req.epochs + 1):
        for g in opt.param_groups:
            g["lr"] = cosine_lr(req.lr, 0.0, req.epochs, e)
        loss = torch.nn.functional.mse_loss(_forward(req, obj, aux, a), a["measured"])
        opt.zero_grad(set_to_none=True)
        loss.backward()
        if e == 1 and first_grads is not None:
            obj.grad.copy_(first_grads[0])
            aux.grad.copy_(first_grads[1])
        opt.step()
        if req.kind == "coded":
            with torch.no_grad():
                aux.copy_(aux.real.clamp_min(0.0).to(torch.complex64))
    return float(loss.detach())


def test_standard_mix_alone_equals_mix(gpu, problems):
    srv = sv.Server(gpu, hbm_budget=16e9, slots=8, keep_results=True)
    mixed = srv.run(MIX, problems)
    assert srv.peak_in_flight > 1
    for req, m in zip(MIX, mixed):
        alone = sv.Server(gpu, hbm_budget=16e9, keep_results=True).run([req], {req.rid: problems[req.rid]})[0]
        assert torch.equal(m.object, alone.object), req
        assert torch.equal(m.aux, alone.aux), req
        assert m.final_loss == alone.final_loss, req


@pytest.mark.parametrize("idx", range(len(MIX)))
def test_standard_mix_class_matches_plain_path(gpu, problems, idx):
    req = MIX[idx]
    prob = problems[req.rid]
    # first epoch: the fused step against the plain path's autograd, same inputs
    job = sv.Job(req, prob, gpu, torch.cuda.current_stream(gpu), graph=False)
    job._fused_step()
    torch.cuda.synchronize()
    a = {k: v.to(gpu) for k, v in prob.arrays.items()}
    o = a["object"].clone().requires_grad_(True)
    x = a["aux"].clone().requires_grad_(True)
    amp = _forward(req, o, x, a)
    torch.nn.functional.mse_loss(amp, a["measured"]).backward()
    per = ((amp.detach() - a["measured"]) ** 2).mean(dim=(1, 2))
    e1 = (nrel(job.step.loss[: job.views], per), nrel(job.g_obj, o.grad), nrel(job.g_aux, x.grad))
    seed = (job.g_obj.clone(), job.g_aux.clone())
    job.release()
    # whole runs
    fused = sv.Server(gpu, hbm_budget=16e9).run([req], {req.rid: prob})[0].final_loss
    plain = _plain(req, prob, gpu)
    seeded = _plain(req, prob, gpu, first_grads=seed)
    gap, seeded_gap = abs(fused - plain) / abs(plain), abs(fused - seeded) / abs(seeded)
    print(f"{req.kind} {req.size} x{req.views}: epoch 1 loss_b {e1[0]:.2e} grad_object {e1[1]:.2e} grad_aux "
          f"{e1[2]:.2e}; final loss fused {fused:.7g} plain {plain:.7g} (rel {gap:.2e}), plain seeded with the "
          f"fused first gradients {seeded:.7g} (rel {seeded_gap:.2e})", flush=True)
    assert max(e1) < 1e-5
    assert seeded_gap < 1e-4
    if req.kind in CLASSES_WITHOUT_EXACT_ZEROS:
        assert gap < 1e-4

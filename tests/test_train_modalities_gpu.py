"""The Fourier, near-field and coded reconstruction loops (fused steps, Adam,
on-device mini-batch gathers, captured epochs) against the reference drivers'
plain torch loops on the same permutations (python/train_fpm.py:229-283,
python/train_coded.py:199-262), and graph replay against eager launches."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.helpers import crandn, disc_probe, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu


def _perms(gpu, P, epochs, seed):
    g = torch.Generator(device=gpu).manual_seed(seed)
    return [torch.randperm(P, generator=g, device=gpu) for _ in range(epochs)]


def _torch_loop(forward, loss_fn, obj, aux, P, perms, batch, lr, aux_lr, epochs, project=False, fix_aux=False):
    """The drivers' loop: per epoch a permutation, mini-batches, loss, backward,
    Adam(betas (0.9, 0.999), eps 1e-16) with cosine learning rates,
    epoch_loss = sum(loss * batch) / views."""
    from ptychox_amd.train import cosine_lr

    o = torch.nn.Parameter(obj.clone())
    a = torch.nn.Parameter(aux.clone()) if not fix_aux else aux.clone()
    groups = [{"params": [o], "lr": lr}] + ([] if fix_aux else [{"params": [a], "lr": aux_lr}])
    opt = torch.optim.Adam(groups, betas=(0.9, 0.999), eps=1e-16)
    out = []
    for epoch in range(1, epochs + 1):
        for g, base in zip(opt.param_groups, (lr, aux_lr)):
            g["lr"] = cosine_lr(base, 0.0, epochs, epoch)
        tot = 0.0
        for start in range(0, P, batch):
            bi = perms[epoch - 1][start:start + batch]
            loss = loss_fn(forward(o, a, bi), bi)
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            if project and not fix_aux:
                with torch.no_grad():
                    a.copy_(a.real.clamp_min(0.0).to(torch.complex64))
            tot += loss.item() * bi.numel()
        out.append(tot / P)
    return out


def _fpm_problem(gpu):
    from ptychox_amd.synthetic import fpm_dataset

    d = fpm_dataset(128, 32, 12.0, 5, 6, gpu, seed=2)
    init = torch.full((128, 128), 0.6, dtype=torch.complex64, device=gpu)
    pupil = (d.pupil.abs() > 0).to(torch.complex64)
    return init, pupil, d.leds, d.measured


@pytest.mark.parametrize("graph", [False, True])
def test_fourier_loop_vs_driver(gpu, graph):
    from ptychox_amd import ptychoX_fourier
    from ptychox_amd.train import FourierReconstruction

    init, pupil, led, meas = _fpm_problem(gpu)
    P, batch, epochs = led.shape[0], 8, 4
    rec = FourierReconstruction(init, pupil, led, meas, batch, lr=2e-2, pupil_lr=1e-2, epochs=epochs,
                                generator=torch.Generator(device=gpu).manual_seed(5), graph=graph)
    res = [r.loss for r in rec.fit()]
    ref = _torch_loop(lambda o, a, bi: ptychoX_fourier(o, a, led[bi]),
                      lambda pred, bi: torch.nn.functional.mse_loss(pred, meas[bi]),
                      init, pupil, P, _perms(gpu, P, epochs, 5), batch, 2e-2, 1e-2, epochs)
    e = nrel(res, ref)
    print(f"fourier loop graph={graph}: {[round(x, 6) for x in res]} vs driver {[round(x, 6) for x in ref]} "
          f"({e:.2e})", flush=True)
    assert res[-1] < res[0] and e < 1e-3


def test_nearfield_loop_vs_driver(gpu):
    from ptychox_amd import ptychoX_nearfield
    from ptychox_amd.synthetic import nearfield_dataset
    from ptychox_amd.train import NearFieldReconstruction

    d = nearfield_dataset(128, 32, 12, 8, 10.0, 1e5, gpu, seed=4)
    init = torch.ones((128, 128), dtype=torch.complex64, device=gpu)
    P, batch, epochs = 12, 4, 4
    rec = NearFieldReconstruction(init, d.probe, d.transfer, d.positions, d.measured, batch, lr=2e-2, epochs=epochs,
                                  generator=torch.Generator(device=gpu).manual_seed(6), graph=True)
    res = [r.loss for r in rec.fit()]
    ref = _torch_loop(lambda o, a, bi: ptychoX_nearfield(o, a, d.transfer, d.positions[bi]),
                      lambda pred, bi: torch.nn.functional.mse_loss(pred, d.measured[bi]),
                      init, d.probe, P, _perms(gpu, P, epochs, 6), batch, 2e-2, 2e-2, epochs)
    e = nrel(res, ref)
    print(f"near-field loop: {[round(x, 6) for x in res]} vs driver {[round(x, 6) for x in ref]} ({e:.2e})",
          flush=True)
    assert res[-1] < res[0] and e < 1e-3


@pytest.mark.parametrize("loss,ds", [("mse", 1), ("mse", 2), ("fd-isotropic", 1), ("fd-anisotropic", 2)])
def test_coded_loop_vs_driver(gpu, loss, ds):
    from ptychox_amd import ptychoX_coded
    from ptychox_amd.losses import fd_loss
    from ptychox_amd.synthetic import coded_dataset
    from ptychox_amd.train import CodedReconstruction

    d = coded_dataset(64, 3, 4, 40.0, 1e4, gpu, seed=8, downsample=ds)
    init_o = torch.ones((64, 64), dtype=torch.complex64, device=gpu)
    init_c = torch.ones((64, 64), dtype=torch.complex64, device=gpu)
    P, batch, epochs = d.shifts.shape[0], 4, 4
    rec = CodedReconstruction(init_o, init_c, d.transfer_1, d.transfer_2, d.shifts, ds, d.measured, batch,
                              lr=2e-2, epochs=epochs, generator=torch.Generator(device=gpu).manual_seed(9),
                              graph=True, loss=loss)
    res = [r.loss for r in rec.fit()]
    assert torch.all(rec.aux.imag == 0) and torch.all(rec.aux.real >= 0)
    if loss == "mse":
        lf = lambda pred, bi: torch.nn.functional.mse_loss(pred, d.measured[bi])
    else:
        lf = lambda pred, bi: fd_loss(pred, d.measured[bi], isotropic=loss == "fd-isotropic")
    ref = _torch_loop(lambda o, a, bi: ptychoX_coded(o, a, d.transfer_1, d.transfer_2, d.shifts[bi], ds), lf,
                      init_o, init_c, P, _perms(gpu, P, epochs, 9), batch, 2e-2, 2e-2, epochs, project=True)
    e = nrel(res, ref)
    print(f"coded loop {loss} ds={ds}: {[round(x, 6) for x in res]} vs driver {[round(x, 6) for x in ref]} "
          f"({e:.2e})", flush=True)
    assert e < 1e-3


def test_coded_loop_graph_equals_eager_and_fixed_surface(gpu):
    from ptychox_amd.synthetic import coded_dataset
    from ptychox_amd.train import CodedReconstruction

    d = coded_dataset(64, 3, 4, 40.0, 1e4, gpu, seed=10)
    init_o = torch.ones((64, 64), dtype=torch.complex64, device=gpu)
    out = []
    for graph in (False, True):
        rec = CodedReconstruction(init_o, d.coded_surface, d.transfer_1, d.transfer_2, d.shifts, 1, d.measured, 4,
                                  lr=2e-2, epochs=3, generator=torch.Generator(device=gpu).manual_seed(1),
                                  graph=graph, loss="fd-isotropic", fix_coded_surface=True)
        losses = [r.loss for r in rec.fit()]
        out.append((losses, rec.obj.clone(), rec.aux.clone()))
    assert out[0][0] == out[1][0]
    assert torch.equal(out[0][1], out[1][1])
    assert torch.equal(out[0][2], d.coded_surface) and torch.equal(out[1][2], d.coded_surface)

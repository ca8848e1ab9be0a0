"""FP64-oracle error of the fused HIP steps at each benchmark configuration's
full size, on that configuration's own synthetic recipe:

* conventional, CPU-runnable case: object 512^2, probe 128^2 (radius 32,
  defocus), 15 x 15 raster (step 24, +-2 px jitter), Poisson 1e4; its first
  mini-batch of 32 and all 225 positions in one call;
* conventional: object 2048^2, probe 512^2 (radius 128, defocus), the first 64
  positions of the 30 x 30 raster (step 52, +-3 px jitter), Poisson 1e5;
* Fourier: spectrum 1024^2, pupil 256^2 (radius 48, random defocus), 15 x 15
  LEDs on a 40 px pitch (225 images), 1 % Gaussian noise;
* near-field: object 1024^2, 512^2 diffuser probe, 64 positions within
  +-32 px, angular spectrum at Fresnel number 20, Poisson 1e4;
* coded: 1024^2, 7 x 7 translations on an 8 px pitch, H1 = H2 = the same
  angular spectrum, Poisson 1e4;
* the five request classes of the served mix (ptychox_amd.serving), each with
  the data and initial guesses a request arrives with: the server's first
  epoch is this fused step.

Each step is evaluated at two estimates against the FP64 oracle's loss and
gradients (oracle/ptycho_oracle.py, restating python/train_ptychography.py:
549-558 and the operators it calls), and so is the reference's own FP32 torch
path on the same GPU:

* the reference drivers' initial guess (python/train_ptychography.py:495-498,
  train_fpm.py:197-204, train_coded.py:172-177): bar 1e-5 norm-wise relative
  for per-position losses, amplitudes and both gradients, and 1e-5 for the
  worst single position's loss (Fourier: as below, the initial guess is a
  smooth real image whose spectrum has the same dynamic range);
* a smoothly perturbed ground truth (a reconstruction near convergence): there
  the residual A - Y is a few percent of A and the gradients cancel, so every
  FP32 evaluation loses digits; the bar is max(1e-5, 2x the reference FP32
  path's error).  The Fourier smooth object also puts the far LEDs'
  sub-spectra ~1e-12 below DC, where every FP32 FFT returns rounding noise.

All errors are printed.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import ptycho_oracle as po
from tests.helpers import nrel

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _perturb(x: torch.Tensor, seed: int, amp: float = 0.05) -> torch.Tensor:
    """x * (1 + amp * smooth complex noise): a nearby smooth estimate."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    n = gaussian_filter(rng.standard_normal(tuple(x.shape)), 4.0, mode="wrap") + 1j * gaussian_filter(
        rng.standard_normal(tuple(x.shape)), 4.0, mode="wrap")
    n = n / np.abs(n).max()
    return (x * torch.from_numpy((1 + amp * n).astype(np.complex64)).to(x.device)).contiguous()


def _uniform_real(shape, seed, device):
    """torch.rand(...).to(complex64) of the conventional driver (its own generator)."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.rand(shape, generator=g).to(torch.complex64).to(device)


def _aperture(x: torch.Tensor) -> torch.Tensor:
    return (x.abs() > 0).to(torch.complex64).contiguous()


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def _per_position(got, ref: np.ndarray) -> float:
    g = _np(got).astype(np.float64)
    return float(np.max(np.abs(g - ref) / np.abs(ref)))


def _errors(per, amp, g1, g2, ref):
    out = {"loss_per_position": nrel(per, ref["per"]), "worst_position": _per_position(per, ref["per"]),
           "grad_1": nrel(g1, ref["g1"]), "grad_2": nrel(g2, ref["g2"])}
    if amp is not None and "amp" in ref:
        out["amplitude"] = nrel(amp, ref["amp"])
    return out


def _torch_fp32(op, leaves, y):
    """The reference's FP32 torch path on the GPU: op -> F.mse_loss -> backward."""
    ls = [x.clone().requires_grad_(True) for x in leaves]
    amp = op(*ls)
    F.mse_loss(amp, y).backward()
    per = ((amp.detach() - y) ** 2).mean(dim=(1, 2))
    return per, amp.detach(), ls[0].grad, ls[1].grad


def _check(name, state, hip, t32, conditioned=True):
    """conditioned: the strict bar applies at the initial guess (not for the
    Fourier smooth spectrum, whose far sub-spectra no FP32 FFT resolves)."""
    fmt = lambda e: ", ".join(f"{k} {v:.2e}" for k, v in e.items())
    print(f"{name} [{state}] vs FP64: HIP {{{fmt(hip)}}}; reference FP32 torch path {{{fmt(t32)}}}", flush=True)
    for k in hip:
        strict = conditioned and state == "initial guess"
        bar = TOL if strict else max(TOL, 2.0 * t32[k])
        assert hip[k] < bar, (name, state, k, hip[k], t32[k])


@pytest.mark.parametrize("B", [32, 225])
def test_conventional_cpu_case_vs_fp64_oracle(gpu, B):
    from ptychox_amd import ptychoX_base
    from ptychox_amd.base import BaseStep
    from ptychox_amd.synthetic import conventional_dataset

    d = conventional_dataset(512, 128, 32.0, 15, 24, 24, 2, 1e4, 4.0, gpu, seed=0)
    pos, y = d.positions[:B].contiguous(), d.measured[:B].contiguous()
    step = BaseStep((512, 512), (128, 128), B, gpu, keep_amplitude=True)
    states = {"initial guess": (_uniform_real((512, 512), 0, gpu), _aperture(d.probe)),
              "near convergence": (_perturb(d.sample, 9), _perturb(d.probe, 10))}
    for state, (s, p) in states.items():
        gs, gp = torch.empty_like(s), torch.empty_like(p)
        lb = step(s, p, pos, y, gs, gp)
        torch.cuda.synchronize()
        sn, pn, posn, yn = _np(s), _np(p), _np(pos), _np(y)
        _, per, gs_ref, gp_ref = po.base_train_step(sn, pn, posn, yn)
        ref = {"per": per, "g1": gs_ref, "g2": gp_ref, "amp": po.base_forward(sn, pn, posn)["amplitude"]}
        hip = _errors(lb, step.amplitude, gs, gp, ref)
        t32 = _errors(*_torch_fp32(lambda a, b: ptychoX_base(a, b, pos), [s, p], y), ref)
        _check(f"conventional 512^2/128^2 B={B}", state, hip, t32)


def test_conventional_full_size_vs_fp64_oracle(gpu):
    from ptychox_amd import ptychoX_base
    from ptychox_amd.base import BaseStep
    from ptychox_amd.synthetic import conventional_dataset

    d = conventional_dataset(2048, 512, 128, 30, 52, 14, 3, 1e5, 4.0, gpu, seed=0)
    B = 64
    pos, y = d.positions[:B].contiguous(), d.measured[:B].contiguous()
    step = BaseStep((2048, 2048), (512, 512), B, gpu, keep_amplitude=True)
    states = {"initial guess": (_uniform_real((2048, 2048), 0, gpu), _aperture(d.probe)),
              "near convergence": (_perturb(d.sample, 1), _perturb(d.probe, 2))}
    for state, (s, p) in states.items():
        gs, gp = torch.empty_like(s), torch.empty_like(p)
        lb = step(s, p, pos, y, gs, gp)
        torch.cuda.synchronize()
        sn, pn, posn, yn = _np(s), _np(p), _np(pos), _np(y)
        _, per, gs_ref, gp_ref = po.base_train_step(sn, pn, posn, yn)
        ref = {"per": per, "g1": gs_ref, "g2": gp_ref, "amp": po.base_forward(sn, pn, posn)["amplitude"]}
        hip = _errors(lb, step.amplitude, gs, gp, ref)
        t32 = _errors(*_torch_fp32(lambda a, b: ptychoX_base(a, b, pos), [s, p], y), ref)
        _check("conventional 2048^2/512^2 B=64", state, hip, t32)


def test_fourier_full_size_vs_fp64_oracle(gpu):
    from ptychox_amd import ptychoX_fourier
    from ptychox_amd.fpm import FpmStep
    from ptychox_amd.synthetic import fpm_dataset

    d = fpm_dataset(1024, 256, 48, 15, 40, gpu, seed=0)
    led, y = d.leds, d.measured
    B = led.shape[0]
    step = FpmStep((1024, 1024), (256, 256), B, gpu, keep_amplitude=True)
    init = F.interpolate(y.mean(dim=0, keepdim=True).unsqueeze(0), size=(1024, 1024), mode="bilinear",
                         align_corners=False)[0, 0].to(torch.complex64).contiguous()
    states = {"initial guess": (init, _aperture(d.pupil)),
              "near convergence": (_perturb(d.object_field, 3), _perturb(d.pupil, 4))}
    for state, (o, q) in states.items():
        go, gq = torch.empty_like(o), torch.empty_like(q)
        lb = step(o, q, led, y, go, gq)
        torch.cuda.synchronize()
        on, qn, ledn, yn = _np(o), _np(q), _np(led), _np(y)
        _, per, go_ref, gq_ref = po.fpm_train_step(on, qn, ledn, yn)
        ref = {"per": per, "g1": go_ref, "g2": gq_ref, "amp": po.fpm_forward(on, qn, ledn)["amplitude"]}
        hip = _errors(lb, step.amplitude, go, gq, ref)
        t32 = _errors(*_torch_fp32(lambda a, b: ptychoX_fourier(a, b, led), [o, q], y), ref)
        _check("fourier 1024^2/256^2 B=225", state, hip, t32, conditioned=False)


def test_nearfield_full_size_vs_fp64_oracle(gpu):
    from ptychox_amd import ptychoX_nearfield
    from ptychox_amd.nearfield import NearFieldStep
    from ptychox_amd.synthetic import nearfield_dataset

    d = nearfield_dataset(1024, 512, 64, 32, 20.0, 1e4, gpu, seed=0)
    t, pos, y = d.transfer, d.positions, d.measured
    B = pos.shape[0]
    step = NearFieldStep((1024, 1024), (512, 512), B, gpu, keep_amplitude=True)
    # no reference driver: the conventional driver's guesses (uniform real object), the known diffuser
    states = {"initial guess": (_uniform_real((1024, 1024), 0, gpu), d.probe),
              "near convergence": (_perturb(d.sample, 5), _perturb(d.probe, 6))}
    for state, (s, p) in states.items():
        gs, gp = torch.empty_like(s), torch.empty_like(p)
        lb = step(s, p, t, pos, y, gs, gp)
        torch.cuda.synchronize()
        sn, pn, tn, posn, yn = _np(s), _np(p), _np(t), _np(pos), _np(y)
        _, per, gs_ref, gp_ref = po.nearfield_train_step(sn, pn, tn, posn, yn)
        ref = {"per": per, "g1": gs_ref, "g2": gp_ref, "amp": po.nearfield_forward(sn, pn, tn, posn)["amplitude"]}
        hip = _errors(lb, step.amplitude, gs, gp, ref)
        t32 = _errors(*_torch_fp32(lambda a, b: ptychoX_nearfield(a, b, t, pos), [s, p], y), ref)
        _check("near-field 1024^2/512^2 B=64", state, hip, t32)


@pytest.mark.timeout(600)
def test_coded_full_size_vs_fp64_oracle(gpu):
    from ptychox_amd import ptychoX_coded
    from ptychox_amd.coded import CodedStep
    from ptychox_amd.synthetic import coded_dataset

    d = coded_dataset(1024, 7, 8, 256.0 ** 2 / 20.0, 1e4, gpu, seed=0)
    t, sh, y = d.transfer_1, d.shifts, d.measured
    B = sh.shape[0]
    step = CodedStep((1024, 1024), B, 1, gpu, keep_amplitude=True)
    ones = torch.ones((1024, 1024), dtype=torch.complex64, device=gpu)
    states = {"initial guess": (ones, ones.clone()),
              "near convergence": (_perturb(d.object_field, 7), _perturb(d.coded_surface, 8))}
    for state, (o, c) in states.items():
        go, gc = torch.empty_like(o), torch.empty_like(c)
        lb = step(o, c, t, t, sh, y, go, gc)
        torch.cuda.synchronize()
        on, cn, tn, shn, yn = _np(o), _np(c), _np(t), _np(sh), _np(y)
        fwd = po.coded_forward(on, cn, tn, tn, shn, 1)
        _, per, g = po.mse_amplitude_loss(fwd["amplitude"], yn)
        go_ref, gc_ref = po.coded_backward(g, fwd, cn, tn, tn, 1)
        ref = {"per": per, "g1": go_ref, "g2": gc_ref, "amp": fwd["amplitude"]}
        del fwd
        hip = _errors(lb, step.amplitude, go, gc, ref)
        t32 = _errors(*_torch_fp32(lambda a, b: ptychoX_coded(a, b, t, t, sh, 1), [o, c], y), ref)
        _check("coded 1024^2 B=49", state, hip, t32)


# one request of each class of the served mix (its first epoch's fused step)
SERVED = [("conventional", 128, 400), ("conventional", 256, 400), ("fourier", 128, 225),
          ("nearfield", 512, 64), ("coded", 512, 25)]


@pytest.mark.parametrize("kind,size,views", SERVED)
def test_served_request_classes_vs_fp64_oracle(gpu, kind, size, views):
    from ptychox_amd import ptychoX_base, ptychoX_coded, ptychoX_fourier, ptychoX_nearfield
    from ptychox_amd import serving as sv
    from ptychox_amd.base import BaseStep
    from ptychox_amd.coded import CodedStep
    from ptychox_amd.fpm import FpmStep
    from ptychox_amd.nearfield import NearFieldStep

    req = sv.Request(0.0, 7, kind, size, views, seed=7)
    a = {k: v.to(gpu) for k, v in sv.make_problem(req).arrays.items()}
    o, q, y = a["object"], a["aux"], a["measured"]
    go, gq = torch.empty_like(o), torch.empty_like(q)
    V = y.shape[0]
    on, qn, yn = _np(o), _np(q), _np(y)
    if kind == "conventional":
        step = BaseStep(o.shape, q.shape, V, gpu, keep_amplitude=True)
        lb = step(o, q, a["positions"], y, go, gq)
        _, per, g1, g2 = po.base_train_step(on, qn, _np(a["positions"]), yn)
        amp = po.base_forward(on, qn, _np(a["positions"]))["amplitude"]
        op = lambda u, v: ptychoX_base(u, v, a["positions"])
    elif kind == "fourier":
        step = FpmStep(o.shape, q.shape, V, gpu, keep_amplitude=True)
        lb = step(o, q, a["positions"], y, go, gq)
        _, per, g1, g2 = po.fpm_train_step(on, qn, _np(a["positions"]), yn)
        amp = po.fpm_forward(on, qn, _np(a["positions"]))["amplitude"]
        op = lambda u, v: ptychoX_fourier(u, v, a["positions"])
    elif kind == "nearfield":
        step = NearFieldStep(o.shape, q.shape, V, gpu, keep_amplitude=True)
        lb = step(o, q, a["transfer"], a["positions"], y, go, gq)
        tn, posn = _np(a["transfer"]), _np(a["positions"])
        _, per, g1, g2 = po.nearfield_train_step(on, qn, tn, posn, yn)
        amp = po.nearfield_forward(on, qn, tn, posn)["amplitude"]
        op = lambda u, v: ptychoX_nearfield(u, v, a["transfer"], a["positions"])
    else:
        step = CodedStep(o.shape, V, 1, gpu, keep_amplitude=True)
        t, sh = a["transfer"], a["shifts"]
        lb = step(o, q, t, t, sh, y, go, gq)
        tn, shn = _np(t), _np(sh)
        fwd = po.coded_forward(on, qn, tn, tn, shn, 1)
        _, per, g = po.mse_amplitude_loss(fwd["amplitude"], yn)
        g1, g2 = po.coded_backward(g, fwd, qn, tn, tn, 1)
        amp = fwd["amplitude"]
        op = lambda u, v: ptychoX_coded(u, v, t, t, sh, 1)
    torch.cuda.synchronize()
    ref = {"per": per, "g1": g1, "g2": g2, "amp": amp}
    hip = _errors(lb, step.amplitude, go, gq, ref)
    t32 = _errors(*_torch_fp32(op, [o, q], y), ref)
    _check(f"served {kind} {size} x{V}", "initial guess", hip, t32)

"""HIP stencil losses against the FP64 oracle (every kind, real residuals and
complex fields, odd sizes that split rows across work-groups), the
reference's fd_loss through torch autograd on the same GPU, determinism, and
the coded training step with --loss fd-isotropic / fd-anisotropic."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import nrel, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5
SHAPES = [(3, 64, 48), (1, 128, 128), (5, 33, 17), (2, 1, 300)]


def _eps(kind):
    return 1e-5 if kind in ("fd_iso", "sparse_iso") else 0.0


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", po.LOSS_KINDS)
def test_amplitude_loss_vs_oracle(gpu, kind, shape):
    from ptychox_amd import _lib
    from ptychox_amd.losses import native_amplitude_loss

    rng = np.random.default_rng(100 * po.LOSS_KINDS.index(kind) + SHAPES.index(shape))
    pred = rng.random(shape).astype(np.float32)
    target = rng.random(shape).astype(np.float32)
    B, h, w = shape
    scale = 0.5
    per = torch.empty(B, device=gpu)
    g = torch.empty(shape, device=gpu)
    ws = _lib.workspace(_lib.query("ptx_loss_workspace", B, h, w), gpu)
    native_amplitude_loss(kind, to_dev(pred, gpu), to_dev(target, gpu), scale, _eps(kind), per, g, ws)
    L, G = po.stencil_loss(kind, pred.astype(np.float64) - target, _eps(kind))
    e = (nrel(per, scale * L), nrel(g, scale * G))
    print(f"{kind} {shape}: loss {e[0]:.2e} grad {e[1]:.2e}", flush=True)
    assert max(e) < TOL


@pytest.mark.parametrize("kind,ty", [("fd_iso", "isotropic"), ("fd_aniso", "anisotropic"), ("grad_l2", "L2")])
def test_tv_regularizer_of_complex_field_vs_oracle(gpu, kind, ty):
    from ptychox_amd.losses import tv_loss

    rng = np.random.default_rng(3)
    z = (rng.standard_normal((96, 80)) + 1j * rng.standard_normal((96, 80))).astype(np.complex64)
    zt = to_dev(z, gpu).requires_grad_(True)
    v = tv_loss(zt, ty, use_cuda_kernels=True)
    v.backward()
    L, G = po.stencil_loss(kind, z.astype(np.complex128), 0.0)
    e = (abs(v.item() - L.sum()) / abs(L.sum()), nrel(zt.grad, G))
    print(f"tv {ty}: loss {e[0]:.2e} grad {e[1]:.2e}", flush=True)
    assert max(e) < TOL


@pytest.mark.parametrize("ty", ["isotropic", "anisotropic"])
def test_sparse_regularizer_vs_oracle(gpu, ty):
    from ptychox_amd.losses import sparse_loss

    rng = np.random.default_rng(4)
    z = (0.05 * (rng.standard_normal((3, 64, 64)) + 1j * rng.standard_normal((3, 64, 64)))).astype(np.complex64)
    zt = to_dev(z, gpu).requires_grad_(True)
    v = sparse_loss(zt, ty, use_cuda_kernels=True)
    v.backward()
    kind = "sparse_iso" if ty == "isotropic" else "sparse_aniso"
    L, G = po.stencil_loss(kind, z.astype(np.complex128), _eps(kind))
    assert abs(v.item() - L.sum()) / abs(L.sum()) < TOL and nrel(zt.grad, G) < TOL


@pytest.mark.parametrize("iso", [True, False])
def test_fd_loss_native_vs_torch_autograd(gpu, iso):
    from ptychox_amd.losses import fd_loss

    rng = np.random.default_rng(5 + iso)
    pred = rng.random((7, 96, 64)).astype(np.float32)
    target = rng.random((7, 96, 64)).astype(np.float32)
    a = to_dev(pred, gpu).requires_grad_(True)
    b = to_dev(pred, gpu).requires_grad_(True)
    t = to_dev(target, gpu)
    ln = fd_loss(a, t, iso, use_cuda_kernels=True)
    lt = fd_loss(b, t, iso, use_cuda_kernels=False)
    ln.backward()
    lt.backward()
    assert abs(ln.item() - lt.item()) / abs(lt.item()) < TOL
    assert nrel(a.grad, b.grad.double().cpu().numpy()) < TOL
    # determinism: the same call twice is bitwise equal
    a2 = to_dev(pred, gpu).requires_grad_(True)
    ln2 = fd_loss(a2, t, iso, use_cuda_kernels=True)
    ln2.backward()
    assert torch.equal(ln, ln2) and torch.equal(a.grad, a2.grad)


@pytest.mark.parametrize("loss", ["fd-isotropic", "fd-anisotropic"])
def test_coded_step_with_fd_loss_vs_torch_driver(gpu, loss):
    # the coded driver's step with --loss fd-*: ptychoX_coded (torch path) +
    # fd_loss + backward, against CodedStep on the HIP path
    from ptychox_amd import ptychoX_coded
    from ptychox_amd.coded import CodedStep
    from ptychox_amd.losses import fd_loss
    from tests.test_coded_gpu import _case

    rng = np.random.default_rng(11)
    n, B, ds = 128, 9, 2
    o, c, t1, t2, sh = _case(rng, n, B)
    fwd = po.coded_forward(o, c, t1, t2, sh.astype(np.float64), ds)
    # residual = a +-0.1 checkerboard plus small noise: every difference dx, dy
    # stays far from 0, so sgn() cannot flip between the two FP32 paths
    a = fwd["amplitude"]
    checker = 0.1 * (-1.0) ** np.add.outer(np.arange(a.shape[1]), np.arange(a.shape[2]))
    meas = (a - checker[None] + 0.01 * rng.standard_normal(a.shape)).astype(np.float32)
    iso = loss == "fd-isotropic"
    ot = to_dev(o, gpu).requires_grad_(True)
    ct = to_dev(c, gpu).requires_grad_(True)
    pred = ptychoX_coded(ot, ct, to_dev(t1, gpu), to_dev(t2, gpu), to_dev(sh, gpu), ds, use_cuda_kernels=False)
    ref = fd_loss(pred, to_dev(meas, gpu), iso)
    ref.backward()
    step = CodedStep((n, n), B, ds, gpu, loss=loss)
    go = torch.empty((n, n), dtype=torch.complex64, device=gpu)
    gc = torch.empty_like(go)
    per = step(to_dev(o, gpu), to_dev(c, gpu), to_dev(t1, gpu), to_dev(t2, gpu), to_dev(sh, gpu), to_dev(meas, gpu),
               go, gc)
    e = (abs(per.sum().item() - ref.item()) / abs(ref.item()), nrel(go, ot.grad.cpu().numpy()),
         nrel(gc, ct.grad.cpu().numpy()))
    print(f"coded step {loss}: loss {e[0]:.2e} grad_object {e[1]:.2e} grad_coded {e[2]:.2e}", flush=True)
    assert max(e) < TOL

"""The MATLAB near-field operator's torch path (CPU) against the FP64 oracle,
its layout adapter, and its argument checks."""

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from ptychox_amd import AdPtychoXNearf, ptychoX_nearfield_matlab


def _case(rng, n, B):
    w1 = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    w2 = np.exp(1j * rng.uniform(0, np.pi, (n, n))) * (rng.random((n, n)) < 0.6)
    f = np.fft.fftfreq(n)
    h2 = np.exp(1j * np.pi * (n / 4) ** 2 / 20.0 * (f[:, None] ** 2 + f[None, :] ** 2))
    sh = rng.uniform(-0.2, 0.2, (B, 2))
    return w1, w2, h2, sh


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.complex64)


@pytest.mark.parametrize("ds", [1, 2, 4])
def test_torch_path_matches_oracle_and_gradient(ds):
    rng = np.random.default_rng(40 + ds)
    n, B = 32, 3
    w1, w2, h2, sh = _case(rng, n, B)
    a = _t(w1).requires_grad_(True)
    b = _t(w2).requires_grad_(True)
    obs = ptychoX_nearfield_matlab(a, b, _t(h2), torch.from_numpy(sh), ds)
    g = rng.standard_normal(obs.shape)
    obs.backward(torch.from_numpy(g).float())
    f = po.nearfield_matlab_forward(w1, w2, h2, sh, ds)
    c = po.coded_forward(w1, w2, np.ones((n, n)), h2, sh, ds)
    e1, e2 = po.coded_backward(g, c, w2, np.ones((n, n)), h2, ds)
    rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    assert rel(obs.detach().numpy(), f["observe"]) < 1e-5
    # PyTorch's gradient convention: the exact adjoint
    assert rel(a.grad.numpy(), e1) < 1e-5 and rel(b.grad.numpy(), e2) < 1e-5


def test_matlab_layout_is_per_frame_transpose():
    rng = np.random.default_rng(44)
    w1, w2, h2, sh = _case(rng, 32, 4)
    s = torch.from_numpy(sh)
    a = ptychoX_nearfield_matlab(_t(w1), _t(w2), _t(h2), s, 2)
    m = ptychoX_nearfield_matlab(_t(w1.T), _t(w2.T), _t(h2.T), s, 2, layout="matlab")
    assert float((m - a.mT).abs().max()) < 1e-5 * float(a.abs().max())


def test_argument_checks():
    z = torch.zeros(32, 32, dtype=torch.complex64)
    s = torch.zeros(2, 2)
    with pytest.raises(ValueError):
        ptychoX_nearfield_matlab(z, z[:16, :16], z, s, 2)
    with pytest.raises(ValueError):
        ptychoX_nearfield_matlab(z, z, z, torch.zeros(2, 3), 2)
    with pytest.raises(ValueError):
        ptychoX_nearfield_matlab(z, z, z, s, 3)
    with pytest.raises(TypeError):
        ptychoX_nearfield_matlab(z, z.real.contiguous(), z, s, 2)
    with pytest.raises(ValueError):
        ptychoX_nearfield_matlab(z, z, z, s, 2, layout="fortran")
    with pytest.raises(ValueError):
        AdPtychoXNearf(large_image_size=96, down_sample_ratio=4, device="cpu")
    op = AdPtychoXNearf(large_image_size=32, batch_size=2, down_sample_ratio=4, device="cpu")
    with pytest.raises(RuntimeError):
        op.backward(torch.zeros(2, 8, 8), z, z, z, s, torch.zeros(2, 8, 8))

"""The MATLAB near-field operator on the HIP path (coded kernels with
transfer_1 = NULL) against the FP64 oracle: forward, PyTorch-convention
gradients, the MATLAB-convention backward (PTX_FLAG_MATLAB_NEARF) and the
column-major layout; norm-wise relative error <= 1e-5 in FP32."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from tests.helpers import crandn, nrel, smooth_object, to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _case(rng, n, B):
    w1 = (smooth_object(rng, n) + 0.05 * crandn(rng, n, n)).astype(np.complex64)
    cells = rng.random((n // 2, n // 2)) < 0.5
    w2 = (np.kron(cells, np.ones((2, 2))) * np.exp(1j * np.kron(rng.uniform(0, np.pi, (n // 2, n // 2)),
                                                                  np.ones((2, 2))))).astype(np.complex64)
    f = np.fft.fftfreq(n)
    h2 = np.exp(1j * np.pi * (n / 4) ** 2 / 20.0 * (f[:, None] ** 2 + f[None, :] ** 2)).astype(np.complex64)
    sh = (rng.uniform(-0.1, 0.1, (B, 2))).astype(np.float32)
    return w1, w2, h2, sh


CASES = [(64, 5, 1), (128, 4, 2), (256, 3, 4)]


@pytest.mark.parametrize("n,B,ds", CASES)
def test_forward_and_torch_gradients_vs_oracle(gpu, n, B, ds):
    from ptychox_amd import ptychoX_nearfield_matlab

    rng = np.random.default_rng(n + B + ds)
    w1, w2, h2, sh = _case(rng, n, B)
    a = to_dev(w1, gpu).requires_grad_(True)
    b = to_dev(w2, gpu).requires_grad_(True)
    obs = ptychoX_nearfield_matlab(a, b, to_dev(h2, gpu), to_dev(sh, gpu), ds, use_cuda_kernels=True)
    g = rng.standard_normal(tuple(obs.shape)).astype(np.float32)
    obs.backward(to_dev(g, gpu))
    sh64 = sh.astype(np.float64)
    f = po.nearfield_matlab_forward(w1, w2, h2, sh64, ds)
    c = po.coded_forward(w1, w2, np.ones((n, n)), h2, sh64, ds)
    e1, e2 = po.coded_backward(g, c, w2, np.ones((n, n)), h2, ds)
    e = (nrel(obs, f["observe"]), nrel(a.grad, e1), nrel(b.grad, e2))
    print(f"matlab near-field {n}/B{B}/ds{ds}: observe {e[0]:.2e} dw1 {e[1]:.2e} dw2 {e[2]:.2e}", flush=True)
    assert max(e) < TOL


@pytest.mark.parametrize("n,B,ds", CASES)
def test_matlab_convention_backward_vs_oracle(gpu, n, B, ds):
    from ptychox_amd import AdPtychoXNearf

    rng = np.random.default_rng(7 * n + B + ds)
    w1, w2, h2, sh = _case(rng, n, B)
    op = AdPtychoXNearf(large_image_size=n, batch_size=B, down_sample_ratio=ds, device=gpu)
    obs = op.forward(to_dev(w1, gpu), to_dev(w2, gpu), to_dev(h2, gpu), to_dev(sh, gpu))
    g = rng.standard_normal(tuple(obs.shape)).astype(np.float32)
    d1, d2, dh, dp = op.backward(to_dev(g, gpu), None, to_dev(w2, gpu), to_dev(h2, gpu), to_dev(sh, gpu), obs)
    assert dh is None and dp is None
    f = po.nearfield_matlab_forward(w1, w2, h2, sh.astype(np.float64), ds)
    r1, r2 = po.nearfield_matlab_backward(g, f, w2, h2, sh.astype(np.float64), ds)
    e = (nrel(obs, f["observe"]), nrel(d1, r1), nrel(d2, r2))
    print(f"matlab near-field backward {n}/B{B}/ds{ds}: observe {e[0]:.2e} dldw1 {e[1]:.2e} dldw2 {e[2]:.2e}",
          flush=True)
    assert max(e) < TOL


def test_column_major_layout(gpu):
    from ptychox_amd import AdPtychoXNearf

    rng = np.random.default_rng(99)
    n, B, ds = 128, 4, 2
    w1, w2, h2, sh = _case(rng, n, B)
    logical = AdPtychoXNearf(n, B, ds, device=gpu, layout="torch")
    colmaj = AdPtychoXNearf(n, B, ds, device=gpu, layout="matlab")
    T = lambda a: to_dev(np.ascontiguousarray(a.T), gpu)
    o1 = logical.forward(to_dev(w1, gpu), to_dev(w2, gpu), to_dev(h2, gpu), to_dev(sh, gpu))
    o2 = colmaj.forward(T(w1), T(w2), T(h2), to_dev(sh, gpu))
    assert nrel(o2, o1.mT.contiguous()) < TOL
    g = rng.standard_normal(tuple(o1.shape)).astype(np.float32)
    a1, a2, _, _ = logical.backward(to_dev(g, gpu), None, to_dev(w2, gpu), to_dev(h2, gpu), to_dev(sh, gpu), o1)
    b1, b2, _, _ = colmaj.backward(to_dev(np.ascontiguousarray(np.swapaxes(g, 1, 2)), gpu), None, T(w2), T(h2),
                                   to_dev(sh, gpu), o2)
    assert nrel(b1, a1.mT.contiguous()) < TOL and nrel(b2, a2.mT.contiguous()) < TOL

"""Pins the FP64 oracle (oracle/ptycho_oracle.py) independently of its own formulas.

Each check is chosen so that a plausible mistake (a dropped N, a missing
conj, a wrong shift direction or index, a transposed x/y) fails at least one:

* brute-force DFT and explicit-loop fftshift on small frames, odd sizes too;
* brute-force crop / wrap-around loops;
* central finite differences of every gradient, Re and Im separately
  (PyTorch convention grad = dL/dRe + i dL/dIm);
* FP64 dot-product adjoint identities of the linear pieces;
* closed-form special cases (identity probe, unit pupil, unit transfers,
  integer coded shifts), Parseval;
* the reference's own python/test_ops.py cases: the same seeds and shapes,
  with torch autograd of the reference's "native" formulas
  (test_ops.py:10-17, 63-70) as the independent gradient.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po

RNG = np.random.default_rng(1234)


def crandn(*shape, rng=RNG):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def nrel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# ---------------------------------------------------------------- FFT pieces

@pytest.mark.parametrize("shape", [(4, 5), (8, 8), (7, 3), (16, 12)])
def test_fft_vs_bruteforce(shape):
    x = crandn(2, *shape)
    assert nrel(np.fft.fft2(x), po.dft2_bruteforce(x)) < 1e-13
    assert nrel(np.fft.ifft2(x), po.dft2_bruteforce(x, inverse=True)) < 1e-13


@pytest.mark.parametrize("n", [4, 5, 8, 9])
def test_fftshift_closed_form(n):
    x = crandn(n, n)
    s = po.fftshift2(x)
    si = po.ifftshift2(x)
    for i in range(n):
        for j in range(n):
            # torch.fft.fftshift = roll by n//2; ifftshift = roll by -(n//2)
            assert s[i, j] == x[(i - n // 2) % n, (j - n // 2) % n]
            assert si[i, j] == x[(i + n // 2) % n, (j + n // 2) % n]
    np.testing.assert_array_equal(po.ifftshift2(s), x)


def test_parseval():
    x = crandn(3, 16, 8)
    assert abs(np.sum(np.abs(np.fft.fft2(x)) ** 2) - 128 * np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(x) ** 2) * 128


# ---------------------------------------------------------------- gather loops

def test_base_crop_bruteforce():
    s = crandn(9, 10)
    pos = np.array([[0, 0], [2, 3], [5, 4], [6, 5]])
    p = po.base_crop(s, pos, 4, 4)
    for b, (x, y) in enumerate(pos):
        for r in range(4):
            for c in range(4):
                assert p[b, r, c] == s[y + r, x + c]  # x is the column


def test_base_crop_rejects_out_of_range():
    with pytest.raises(ValueError):
        po.base_crop(crandn(8, 8), np.array([[6, 0]]), 4, 4)


def test_fpm_indices_bruteforce():
    oh, ow, lh, lw = 16, 12, 5, 4
    led = np.array([[0, 0], [11, 2], [7, 15], [-3, 9]])
    idx = po.fpm_flat_indices((oh, ow), (lh, lw), led)
    for b, (x, y) in enumerate(led):
        for r in range(lh):
            for c in range(lw):
                yy = (y + r + oh // 2) % oh
                xx = (x + c + ow // 2) % ow
                assert idx[b, r, c] == yy * ow + xx


# ---------------------------------------------------------------- closed forms

def test_base_closed_form_full_window():
    o = crandn(8, 8)
    out = po.base_forward(o, np.ones((8, 8)), np.array([[0, 0]]))
    np.testing.assert_allclose(out["amplitude"][0], np.abs(np.fft.fftshift(po.dft2_bruteforce(o))), atol=1e-12)


def test_fpm_closed_form_identity():
    o = crandn(8, 8)
    out = po.fpm_forward(o, np.ones((8, 8)), np.array([[0, 0]]))
    np.testing.assert_allclose(out["amplitude"][0], np.abs(o), atol=1e-12)


def test_coded_closed_form_identity_and_integer_shift():
    o = crandn(8, 16)
    one = np.ones((8, 16))
    out = po.coded_forward(o, one, one, one, np.array([[0.0, 0.0]]), 1)
    np.testing.assert_allclose(out["amplitude"][0], np.abs(o), atol=1e-12)
    # shift dx = t / W translates by -t columns; dy = u / H by -u rows
    t, u = 3, 2
    out = po.coded_forward(o, one, one, one, np.array([[t / 16, u / 8]]), 1)
    np.testing.assert_allclose(out["amplitude"][0], np.abs(np.roll(o, (-u, -t), axis=(0, 1))), atol=1e-12)
    # downsample=2 is the RMS over 2x2 cells
    out2 = po.coded_forward(o, one, one, one, np.array([[0.0, 0.0]]), 2)
    a2 = np.abs(o) ** 2
    rms = np.sqrt(a2.reshape(4, 2, 8, 2).mean(axis=(1, 3)))
    np.testing.assert_allclose(out2["amplitude"][0], rms, atol=1e-12)


def test_nearfield_closed_form_unit_transfer():
    s = crandn(12, 12)
    p = crandn(4, 4)
    pos = np.array([[1, 2], [5, 0]])
    out = po.nearfield_forward(s, p, np.ones((4, 4)), pos)
    np.testing.assert_allclose(out["amplitude"], np.abs(po.base_crop(s, pos, 4, 4) * p[None]), atol=1e-12)


# ---------------------------------------------------------------- adjoint identities

def test_adjoint_identities():
    """<A x, y> = <x, A^H y> for the linear pieces the adjoints are made of."""
    x = crandn(3, 8, 4)
    y = crandn(3, 8, 4)
    n = 32
    dot = lambda a, b: np.vdot(b, a)  # <a, b> = sum a conj(b)
    # F^H = N ifft2
    assert abs(dot(np.fft.fft2(x), y) - dot(x, n * np.fft.ifft2(y))) < 1e-10
    # (F^-1)^H = fft2 / N
    assert abs(dot(np.fft.ifft2(x), y) - dot(x, np.fft.fft2(y) / n)) < 1e-10
    # fftshift^H = ifftshift
    assert abs(dot(po.fftshift2(x), y) - dot(x, po.ifftshift2(y))) < 1e-10
    # crop^H = scatter-add
    s = crandn(10, 9)
    pos = np.array([[0, 1], [3, 2], [5, 6]])
    cy = crandn(3, 4, 4)
    adj = np.zeros_like(s)
    for b, (px, py) in enumerate(pos):
        adj[py : py + 4, px : px + 4] += cy[b]
    assert abs(dot(po.base_crop(s, pos, 4, 4), cy) - dot(s, adj)) < 1e-10
    # FPM gather^H = index-add (wrap)
    led = np.array([[0, 0], [9, 3], [-2, 7]])
    idx = po.fpm_flat_indices((10, 9), (4, 4), led)
    acc = np.zeros(90, dtype=complex)
    np.add.at(acc, idx.ravel(), cy.ravel())
    assert abs(dot(s.ravel()[idx], cy) - dot(s.ravel(), acc)) < 1e-10


# ---------------------------------------------------------------- finite differences

def _fd_check(loss_fn, z, grad, entries, h=1e-3, rtol=3e-5):
    """Central differences on Re and Im of ``entries`` of complex array ``z``.

    With h=1e-3 the O(h^2) truncation is ~1e-6 relative; a wrong factor, sign
    or conjugate shows up as an O(1) relative error."""
    for idx in entries:
        for unit, part in ((1.0, np.real), (1j, np.imag)):
            zp = z.copy()
            zm = z.copy()
            zp[idx] += unit * h
            zm[idx] -= unit * h
            fd = (loss_fn(zp) - loss_fn(zm)) / (2 * h)
            an = part(grad[idx])
            assert abs(fd - an) <= rtol * max(1.0, abs(an)) + 1e-8, (idx, part.__name__, fd, an)


def _pick(shape, k, rng):
    return [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(k)]


def test_base_gradient_fd():
    rng = np.random.default_rng(5)
    s = crandn(12, 13, rng=rng)
    p = crandn(8, 8, rng=rng)
    pos = np.array([[0, 0], [3, 2], [5, 4], [4, 4]])
    g = rng.standard_normal((4, 8, 8))
    L = lambda s_, p_: float(np.sum(g * po.base_forward(s_, p_, pos)["amplitude"]))
    fwd = po.base_forward(s, p, pos)
    gs, gp = po.base_backward(g, s.shape, p, fwd["patches"], fwd["far_field"], pos)
    _fd_check(lambda z: L(z, p), s, gs, _pick(s.shape, 12, rng))
    _fd_check(lambda z: L(s, z), p, gp, _pick(p.shape, 12, rng))


def test_fpm_gradient_fd():
    rng = np.random.default_rng(6)
    o = crandn(16, 16, rng=rng)
    q = crandn(8, 8, rng=rng)
    led = np.array([[0, 0], [2, 1], [6, 5], [-3, 4]])
    g = rng.standard_normal((4, 8, 8))
    L = lambda o_, q_: float(np.sum(g * po.fpm_forward(o_, q_, led)["amplitude"]))
    fwd = po.fpm_forward(o, q, led)
    go, gq = po.fpm_backward(g, o.shape, q, fwd["sub_spectrum"], fwd["image_field"], led)
    _fd_check(lambda z: L(z, q), o, go, _pick(o.shape, 12, rng))
    _fd_check(lambda z: L(o, z), q, gq, _pick(q.shape, 12, rng))


@pytest.mark.parametrize("ds", [1, 2])
def test_coded_gradient_fd(ds):
    rng = np.random.default_rng(7 + ds)
    o = crandn(16, 16, rng=rng)
    c = crandn(16, 16, rng=rng)
    t1 = crandn(16, 16, rng=rng)
    t2 = crandn(16, 16, rng=rng)
    sh = np.array([[0.01, -0.02], [0.03, 0.01], [-0.11, 0.07]])
    g = rng.standard_normal((3, 16 // ds, 16 // ds))
    L = lambda o_, c_: float(np.sum(g * po.coded_forward(o_, c_, t1, t2, sh, ds)["amplitude"]))
    fwd = po.coded_forward(o, c, t1, t2, sh, ds)
    go, gc = po.coded_backward(g, fwd, c, t1, t2, ds)
    _fd_check(lambda z: L(z, c), o, go, _pick(o.shape, 12, rng))
    _fd_check(lambda z: L(o, z), c, gc, _pick(c.shape, 12, rng))


def test_nearfield_gradient_fd():
    rng = np.random.default_rng(9)
    s = crandn(14, 12, rng=rng)
    p = crandn(8, 8, rng=rng)
    t = np.exp(1j * rng.uniform(0, 2 * np.pi, (8, 8)))
    pos = np.array([[0, 0], [3, 5], [4, 6]])
    g = rng.standard_normal((3, 8, 8))
    L = lambda s_, p_: float(np.sum(g * po.nearfield_forward(s_, p_, t, pos)["amplitude"]))
    fwd = po.nearfield_forward(s, p, t, pos)
    gs, gp = po.nearfield_backward(g, s.shape, p, t, fwd["patches"], fwd["field"], pos)
    _fd_check(lambda z: L(z, p), s, gs, _pick(s.shape, 12, rng))
    _fd_check(lambda z: L(s, z), p, gp, _pick(p.shape, 12, rng))


def test_mse_loss_gradient_fd():
    rng = np.random.default_rng(11)
    a = rng.uniform(0, 2, (3, 4, 4))
    y = rng.uniform(0, 2, (3, 4, 4))
    loss, per, g = po.mse_amplitude_loss(a, y)
    assert abs(loss - np.mean(per)) < 1e-14
    for idx in _pick(a.shape, 6, rng):
        ap, am = a.copy(), a.copy()
        ap[idx] += 1e-4
        am[idx] -= 1e-4
        fd = (po.mse_amplitude_loss(ap, y)[0] - po.mse_amplitude_loss(am, y)[0]) / 2e-4
        assert abs(fd - g[idx]) < 1e-9


def test_train_step_matches_composition():
    rng = np.random.default_rng(12)
    s = crandn(20, 20, rng=rng)
    p = crandn(8, 8, rng=rng)
    pos = np.array([[0, 0], [5, 7], [12, 3]])
    y = rng.uniform(0, 3, (3, 8, 8))
    loss, per, gs, gp = po.base_train_step(s, p, pos, y)
    # gradient of the loss by finite differences through the whole step
    L = lambda z: po.base_train_step(z, p, pos, y)[0]
    _fd_check(L, s, gs, _pick(s.shape, 6, rng), h=1e-5)


# ---------------------------------------------------------------- reference test_ops.py cases

def _t2n(t):
    return t.detach().to(torch.complex128).numpy() if t.is_complex() else t.detach().double().numpy()


def test_reference_test_ops_base_case():
    """python/test_ops.py:26-49 with the reference's native formula (10-17)."""
    torch.manual_seed(0)
    sample0 = torch.randn(9, 10, dtype=torch.complex64)
    probe0 = torch.randn(4, 5, dtype=torch.complex64)
    positions = torch.tensor([[0, 0], [2, 3], [5, 4]], dtype=torch.long)
    grad = torch.randn(3, 4, 5)
    s = sample0.to(torch.complex128).requires_grad_(True)
    p = probe0.to(torch.complex128).requires_grad_(True)
    patches = torch.stack([s[y : y + 4, x : x + 5] for x, y in positions.tolist()])
    out = torch.fft.fftshift(torch.fft.fft2(patches * p[None]), dim=(-2, -1)).abs()
    (out * grad.double()).sum().backward()
    fwd = po.base_forward(_t2n(sample0), _t2n(probe0), positions.numpy())
    gs, gp = po.base_backward(_t2n(grad), (9, 10), _t2n(probe0), fwd["patches"], fwd["far_field"], positions.numpy())
    assert nrel(fwd["amplitude"], _t2n(out)) < 1e-12
    assert nrel(gs, _t2n(s.grad)) < 1e-12
    assert nrel(gp, _t2n(p.grad)) < 1e-12


def test_reference_test_ops_fpm_case():
    """python/test_ops.py:51-75 (explicit modular-index reference, 63-70)."""
    torch.manual_seed(0)
    torch.randn(9, 10, dtype=torch.complex64)
    torch.randn(4, 5, dtype=torch.complex64)
    torch.randn(3, 4, 5)
    object0 = torch.randn(16, 16, dtype=torch.complex64)
    pupil0 = torch.randn(8, 8, dtype=torch.complex64)
    leds = torch.tensor([[0, 0], [2, 1], [6, 5]], dtype=torch.long)
    fgrad = torch.randn(3, 8, 8)
    o = object0.to(torch.complex128).requires_grad_(True)
    q = pupil0.to(torch.complex128).requires_grad_(True)
    offt = torch.fft.fft2(o)
    patches = []
    for x0, y0 in leds.tolist():
        rows = (torch.arange(8) + y0 + 8) % 16
        cols = (torch.arange(8) + x0 + 8) % 16
        patches.append(offt[rows[:, None], cols[None, :]])
    sub = torch.stack(patches, dim=0)
    ref = torch.fft.ifft2(torch.fft.fftshift(sub * q[None], dim=(-2, -1))).abs()
    (ref * fgrad.double()).sum().backward()
    fwd = po.fpm_forward(_t2n(object0), _t2n(pupil0), leds.numpy())
    go, gq = po.fpm_backward(_t2n(fgrad), (16, 16), _t2n(pupil0), fwd["sub_spectrum"], fwd["image_field"], leds.numpy())
    assert nrel(fwd["amplitude"], _t2n(ref)) < 1e-12
    assert nrel(go, _t2n(o.grad)) < 1e-12
    assert nrel(gq, _t2n(q.grad)) < 1e-12


def test_reference_test_ops_coded_case():
    """python/test_ops.py:77-104: coded op under autograd (coded/ops.py:136-154)."""
    torch.manual_seed(0)
    for shp in ((9, 10), (4, 5)):
        torch.randn(*shp, dtype=torch.complex64)
    torch.randn(3, 4, 5)
    torch.randn(16, 16, dtype=torch.complex64)
    torch.randn(8, 8, dtype=torch.complex64)
    torch.randn(3, 8, 8)
    o0 = torch.randn(16, 16, dtype=torch.complex64)
    c0 = torch.randn(16, 16, dtype=torch.complex64)
    t1 = torch.randn(16, 16, dtype=torch.complex64)
    t2 = torch.randn(16, 16, dtype=torch.complex64)
    shifts = torch.tensor([[0.01, -0.02], [0.03, 0.01]], dtype=torch.float32)
    cgrad = torch.randn(2, 8, 8)
    o = o0.to(torch.complex128).requires_grad_(True)
    c = c0.to(torch.complex128).requires_grad_(True)
    fy = torch.fft.fftfreq(16, dtype=torch.float64) * 16
    yy, xx = torch.meshgrid(fy, fy, indexing="ij")
    sh = shifts.double()
    ramp = torch.exp(1j * 2 * np.pi * (sh[:, 0, None, None] * xx[None] + sh[:, 1, None, None] * yy[None]))
    xf = torch.fft.ifft2(torch.fft.fft2(o)[None] * t1.to(torch.complex128)[None] * ramp)
    z = torch.fft.ifft2(torch.fft.fft2(xf * c[None]) * t2.to(torch.complex128)[None])
    pooled = torch.nn.functional.avg_pool2d(z.abs().square().unsqueeze(1), 2, 2)[:, 0]
    amp = torch.sqrt(pooled.clamp_min(0.0))
    (amp * cgrad.double()).sum().backward()
    fwd = po.coded_forward(_t2n(o0), _t2n(c0), _t2n(t1), _t2n(t2), shifts.double().numpy(), 2)
    go, gc = po.coded_backward(_t2n(cgrad), fwd, _t2n(c0), _t2n(t1), _t2n(t2), 2)
    assert nrel(fwd["amplitude"], _t2n(amp)) < 1e-12
    assert nrel(go, _t2n(o.grad)) < 1e-12
    assert nrel(gc, _t2n(c.grad)) < 1e-12


# ---------------------------------------------------------------- literal FP32 CPU path

def test_torch_literal_path_matches_oracle():
    from oracle import torch_literal as tl

    rng = np.random.default_rng(3)
    s = crandn(40, 40, rng=rng)
    p = crandn(16, 16, rng=rng)
    pos = np.array([[0, 0], [5, 9], [20, 3], [24, 24]])
    y = rng.uniform(0, 5, (4, 16, 16))
    loss, _, gs, gp = po.base_train_step(s, p, pos, y)
    l32, gs32, gp32 = tl.base_step(
        torch.from_numpy(s).to(torch.complex64),
        torch.from_numpy(p).to(torch.complex64),
        torch.from_numpy(pos),
        torch.from_numpy(y).float(),
    )
    assert abs(float(l32) - loss) / loss < 1e-5
    assert nrel(gs32.numpy(), gs) < 1e-5
    assert nrel(gp32.numpy(), gp) < 1e-5


# ---------------------------------------------------------------- MATLAB near-field operator

def test_nearfield_matlab_forward_closed_form_and_coded_form():
    rng = np.random.default_rng(21)
    n = 16
    w1 = crandn(n, n, rng=rng)
    # unit diffuser and propagator, integer shifts t: observe = |roll(w1, -t)|
    t = np.array([[3, -2], [0, 5], [-7, 1]])
    obs = po.nearfield_matlab_forward(w1, np.ones((n, n)), np.ones((n, n)), t / n, 1)["observe"]
    for b in range(3):
        expect = np.abs(np.roll(w1, shift=(-t[b, 1], -t[b, 0]), axis=(0, 1)))
        assert np.allclose(obs[b], expect, atol=1e-12)
    # general case: the coded forward with transfer_1 = 1 (fwd_adPtychoX_nearf.cu:63-114)
    w2, h2 = crandn(n, n, rng=rng), np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n)))
    sh = rng.uniform(-0.3, 0.3, (4, 2))
    for ds in (1, 2, 4):
        a = po.nearfield_matlab_forward(w1, w2, h2, sh, ds)["observe"]
        b = po.coded_forward(w1, w2, np.ones((n, n)), h2, sh, ds)["amplitude"]
        assert nrel(a, b) < 1e-13


@pytest.mark.parametrize("ds", [1, 2])
def test_nearfield_matlab_backward_is_conjugated_adjoint(ds):
    # the MATLAB backward equals the conjugate of the (finite-difference pinned)
    # coded adjoint with H1 = 1, once its detector step Z G/(A + 0.001) is
    # written as the exact one's Z G'/(ds^2 max(A, eps))
    rng = np.random.default_rng(22 + ds)
    n = 16
    w1, w2 = crandn(n, n, rng=rng), crandn(n, n, rng=rng)
    h2 = np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n)))
    sh = rng.uniform(-0.3, 0.3, (3, 2))
    f = po.nearfield_matlab_forward(w1, w2, h2, sh, ds)
    g = rng.standard_normal(f["observe"].shape)
    d1, d2 = po.nearfield_matlab_backward(g, f, w2, h2, sh, ds)
    c = po.coded_forward(w1, w2, np.ones((n, n)), h2, sh, ds)
    a = c["amplitude"]
    g_exact = g * ds * ds * np.maximum(a, po.CODED_AMP_EPS) / (a + po.MATLAB_AMP_EPS)
    e1, e2 = po.coded_backward(g_exact, c, w2, np.ones((n, n)), h2, ds)
    assert nrel(d1, np.conj(e1)) < 1e-13
    assert nrel(d2, np.conj(e2)) < 1e-13
    # and it is not the exact gradient: the detector step and conjugation differ
    x1, _ = po.coded_backward(g, c, w2, np.ones((n, n)), h2, ds)
    assert nrel(d1, x1) > 1e-3


def test_nearfield_matlab_column_major_layout():
    # column-major memory viewed row-major is the per-frame transpose: the same
    # operator with dx and dy swapped returns the transposed frames
    rng = np.random.default_rng(25)
    n = 16
    w1, w2 = crandn(n, n, rng=rng), crandn(n, n, rng=rng)
    h2 = np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n)))
    sh = rng.uniform(-0.3, 0.3, (3, 2))
    f = po.nearfield_matlab_forward(w1, w2, h2, sh, 2)
    ft = po.nearfield_matlab_forward(w1.T, w2.T, h2.T, sh[:, ::-1], 2)
    assert nrel(ft["observe"], np.swapaxes(f["observe"], 1, 2)) < 1e-13
    g = rng.standard_normal(f["observe"].shape)
    d1, d2 = po.nearfield_matlab_backward(g, f, w2, h2, sh, 2)
    t1, t2 = po.nearfield_matlab_backward(np.swapaxes(g, 1, 2), ft, w2.T, h2.T, sh[:, ::-1], 2)
    assert nrel(t1, d1.T) < 1e-13 and nrel(t2, d2.T) < 1e-13


# ---------------------------------------------------------------- alternative losses

def _loop_penalty(kind, d, eps):
    # explicit per-site loops with modular neighbours (no roll)
    B, h, w = d.shape
    out = np.zeros(B)
    for b in range(B):
        for i in range(h):
            for j in range(w):
                dx = d[b, i, (j + 1) % w] - d[b, i, j]
                dy = d[b, (i + 1) % h, j] - d[b, i, j]
                if kind == "l2":
                    v = abs(d[b, i, j]) ** 2
                elif kind == "fd_iso":
                    v = np.sqrt(abs(dx) ** 2 + abs(dy) ** 2 + eps)
                elif kind == "fd_aniso":
                    v = abs(dx) + abs(dy)
                elif kind == "grad_l2":
                    v = abs(dx) ** 2 + abs(dy) ** 2
                elif kind == "sparse_iso":
                    v = 1 - np.exp(-8 * np.sqrt(abs(dx) ** 2 + abs(dy) ** 2 + eps))
                else:
                    v = 1 - np.exp(-8 * (abs(dx) + abs(dy)))
                out[b] += v
    return out


@pytest.mark.parametrize("kind", po.LOSS_KINDS)
@pytest.mark.parametrize("cplx", [False, True])
def test_stencil_loss_values_and_fd_gradient(kind, cplx):
    rng = np.random.default_rng(31 + po.LOSS_KINDS.index(kind) + 7 * cplx)
    d = rng.standard_normal((2, 5, 6)) * 0.3
    if cplx:
        d = d + 1j * rng.standard_normal((2, 5, 6)) * 0.3
    eps = 1e-5 if kind in ("fd_iso", "sparse_iso") else 0.0
    L, g = po.stencil_loss(kind, d, eps)
    assert np.allclose(L, _loop_penalty(kind, d, eps), rtol=1e-12, atol=1e-12)
    f = lambda z: float(np.sum(po.stencil_loss(kind, z, eps)[0]))
    # real data are checked as complex points: the imaginary partials are 0 there
    # (a small step: the square-root penalties curve sharply where |dx|, |dy| are small)
    _fd_check(f, d.astype(np.complex128), g.astype(np.complex128), _pick(d.shape, 14, rng), h=1e-6, rtol=1e-6)


def test_fd_loss_matches_torch_autograd_of_the_reference_formula():
    # train_coded.py:101-107 evaluated with torch autograd, FP64
    rng = np.random.default_rng(35)
    pred, target = rng.random((3, 8, 10)), rng.random((3, 8, 10))
    for iso, kind in ((True, "fd_iso"), (False, "fd_aniso")):
        p = torch.tensor(pred, requires_grad=True)
        diff = p - torch.tensor(target)
        dx = torch.roll(diff, shifts=-1, dims=-1) - diff
        dy = torch.roll(diff, shifts=-1, dims=-2) - diff
        loss = torch.sqrt(dx.square() + dy.square() + 1e-5).sum() if iso else (dx.abs() + dy.abs()).sum()
        loss.backward()
        L, g = po.stencil_loss(kind, pred - target, 1e-5 if iso else 0.0)
        assert abs(L.sum() - loss.item()) < 1e-10 * abs(loss.item())
        assert np.allclose(g, p.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_hessian_loss_loops():
    rng = np.random.default_rng(36)
    w = rng.standard_normal((4, 5)) + 1j * rng.standard_normal((4, 5))
    h_, w_ = w.shape
    x = lambda i, j: w[i % h_, j % w_]
    iso = an = l2 = 0.0
    for i in range(h_):
        for j in range(w_):
            dx = lambda a, b: x(a, b + 1) - x(a, b)
            dy = lambda a, b: x(a + 1, b) - x(a, b)
            dxx = dx(i, j + 1) - dx(i, j)
            dyy = dy(i + 1, j) - dy(i, j)
            dxy = dx(i + 1, j) - dx(i, j)
            iso += np.sqrt(abs(dxx) ** 2 + abs(dyy) ** 2 + 2 * abs(dxy) ** 2)
            an += abs(dxx) + abs(dyy) + 2 * abs(dxy)
            l2 += abs(dx(i, j)) ** 2 + abs(dy(i, j)) ** 2
    assert np.isclose(po.hessian_loss(w, "isotropic"), iso)
    assert np.isclose(po.hessian_loss(w, "anisotropic"), an)
    assert np.isclose(po.hessian_loss(w, "L2"), l2)

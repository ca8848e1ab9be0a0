"""The MATLAB near-field operator on the HIP path.

The reference ships full-frame near-field ptychography as a MATLAB class,
``adPtychoX_nearf`` (matlab/PtychoX_updated/@adPtychoX_nearf/adPtychoX_nearf.m),
over two MEX kernels:

* forward (scripts/forward/fwd_adPtychoX_nearf.cu:58-114):
  ``W1 = fft2(wavefront1)``; ``X_b = ifft2(W1 exp(+i 2 pi (kx dx_b + ky dy_b)))``
  (signed integer frequencies, kernels_nearf.cuh:34-80);
  ``Z_b = ifft2(fft2(X_b wavefront2) H_d2)``;
  ``observe_b = sqrt(mean over ds x ds cells of |Z_b|^2)`` -- the coded forward
  of ``ptychoX_coded`` with ``transfer_1 = 1``;
* backward (scripts/backward/bwd_adPtychoX_nearf.cu): the MATLAB gradient
  convention -- detector step ``Z G/(observe + 0.001)`` without the ``1/ds^2``
  of the exact derivative (kernels_nearf.cuh:507-530) and both gradients
  conjugated relative to PyTorch's ``dL/dRe + i dL/dIm``.

:class:`AdPtychoXNearf` mirrors the class (``forward(wavefront1, wavefront2,
H_d2, shiftsPos)``, ``backward(dl_dout, wavefront1, wavefront2, H_d2,
shiftsPos, observe)`` returning ``(dldw1, dldw2, dldH2, dldPos)`` with the last
two empty, the MATLAB convention).  :func:`ptychoX_nearfield_matlab` is the same
forward as a differentiable torch op with PyTorch's gradient convention (the
exact adjoint), so it drops into a torch training loop.

Layouts: ``layout="torch"`` takes arrays in logical (row, column) order, as
``scipy.io.loadmat`` returns them.  ``layout="matlab"`` takes the raw
column-major memory of MATLAB arrays viewed as row-major tensors (each frame
the transpose of the logical one, as a MEX / gpuArray pointer hands them over);
results come back in the same layout.  The transform commutes with
transposition, so this only swaps the roles of ``dx`` and ``dy``.
"""

from __future__ import annotations

import torch

from . import _lib
from ._utils import require_native_frame
from .coded import native_backward, native_forward, shift_phase_ramp

_LAYOUTS = ("torch", "matlab")


def _check(wavefront1, wavefront2, h_d2, shifts, ds):
    if wavefront1.ndim != 2 or wavefront1.shape[0] != wavefront1.shape[1]:
        raise ValueError("wavefront1 must be a square 2D complex tensor.")
    for name, t in (("wavefront2", wavefront2), ("H_d2", h_d2)):
        if t.shape != wavefront1.shape:
            raise ValueError(f"{name} must have the same shape as wavefront1.")
        if t.device != wavefront1.device:
            raise ValueError(f"{name} must live on the same device as wavefront1.")
        if not t.is_complex():
            raise TypeError(f"{name} must be a complex tensor.")
    if not wavefront1.is_complex():
        raise TypeError("wavefront1 must be a complex tensor.")
    if shifts.ndim != 2 or shifts.shape[-1] != 2:
        raise ValueError("shiftsPos must have shape [scans, 2].")
    if ds < 1 or wavefront1.shape[0] % ds != 0:
        raise ValueError("the frame side must be divisible by down_sample_ratio.")


def _shifts_for(shifts: torch.Tensor, layout: str, device) -> torch.Tensor:
    if layout not in _LAYOUTS:
        raise ValueError(f"layout must be one of {_LAYOUTS}")
    s = shifts.to(device=device, dtype=torch.float32)
    # column-major memory: each frame is the transpose, so dx and dy swap roles
    return (s.flip(-1) if layout == "matlab" else s).contiguous()


def _torch_forward(w1, w2, h2, shifts, ds):
    h, w = w1.shape
    ramp = shift_phase_ramp(h, w, shifts, w1.device, w1.real.dtype)
    x = torch.fft.ifft2(torch.fft.fft2(w1).unsqueeze(0) * ramp)
    z = torch.fft.ifft2(torch.fft.fft2(x * w2.unsqueeze(0)) * h2.unsqueeze(0))
    pooled = torch.nn.functional.avg_pool2d(z.abs().square().unsqueeze(1), kernel_size=ds, stride=ds)
    return torch.sqrt(pooled[:, 0].clamp_min(0.0))


class _NearfieldMatlabFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w1, w2, h2, shifts, ds):  # type: ignore[override]
        w1, w2, h2 = (t.to(torch.complex64).contiguous() for t in (w1, w2, h2))
        spec, x_fwd, sensor, amp = native_forward(w1, w2, None, h2, shifts, ds)
        ctx.save_for_backward(spec, x_fwd, sensor, w2, h2, shifts)
        ctx.ds = int(ds)
        return amp

    @staticmethod
    def backward(ctx, grad_output):  # type: ignore[override]
        spec, x_fwd, sensor, w2, h2, shifts = ctx.saved_tensors
        go, gc = native_backward(grad_output.to(torch.float32).contiguous(), spec, x_fwd, sensor, w2, None, h2,
                                 shifts, ctx.ds)
        return go, gc, None, None, None


def ptychoX_nearfield_matlab(
    wavefront1: torch.Tensor,
    wavefront2: torch.Tensor,
    H_d2: torch.Tensor,
    shiftsPos: torch.Tensor,
    down_sample_ratio: int = 4,
    layout: str = "torch",
    use_cuda_kernels: bool = False,
) -> torch.Tensor:
    """``observe [scans, n/ds, n/ds]`` of the MATLAB near-field operator,
    differentiable w.r.t. ``wavefront1`` and ``wavefront2`` with PyTorch's
    gradient convention (``H_d2`` and ``shiftsPos`` get none, as in the MATLAB
    class).  ``use_cuda_kernels=True`` on complex64 GPU tensors runs the HIP
    coded path with ``transfer_1 = NULL``; otherwise unfused torch ops."""
    ds = int(down_sample_ratio)
    _check(wavefront1, wavefront2, H_d2, shiftsPos, ds)
    shifts = _shifts_for(shiftsPos, layout, wavefront1.device)
    native = (use_cuda_kernels and wavefront1.is_cuda and wavefront1.dtype == torch.complex64
              and wavefront2.dtype == torch.complex64)
    if native:
        require_native_frame(*wavefront1.shape, "wavefront1")
        return _NearfieldMatlabFunction.apply(wavefront1, wavefront2, H_d2, shifts, ds)
    return _torch_forward(wavefront1, wavefront2, H_d2, shifts, ds)


class AdPtychoXNearf:
    """The MATLAB class ``adPtychoX_nearf`` on the HIP path: ``forward`` keeps
    the fields the backward needs (its ``Xfwd_buffer`` / ``Xrcd_buffer``), and
    ``backward`` returns the MATLAB gradient convention."""

    def __init__(self, large_image_size: int = 512, batch_size: int = 16, down_sample_ratio: int = 4,
                 device="cuda", layout: str = "torch"):
        self.n = int(large_image_size)
        require_native_frame(self.n, self.n, "large_image_size")
        self.batch_size = int(batch_size)
        self.mag = int(down_sample_ratio)
        if self.n % self.mag != 0:
            raise ValueError("large_image_size must be divisible by down_sample_ratio.")
        if layout not in _LAYOUTS:
            raise ValueError(f"layout must be one of {_LAYOUTS}")
        self.layout = layout
        self.device = torch.device(device)
        self._saved = None

    def forward(self, wavefront1, wavefront2, H_d2, shiftsPos) -> torch.Tensor:
        _check(wavefront1, wavefront2, H_d2, shiftsPos, self.mag)
        if wavefront1.shape != (self.n, self.n):
            raise ValueError(f"wavefront1 must be {self.n} x {self.n} (large_image_size)")
        if shiftsPos.shape[0] > self.batch_size:
            raise ValueError(f"{shiftsPos.shape[0]} scans exceed batch_size {self.batch_size}")
        w1, w2, h2 = (t.to(device=self.device, dtype=torch.complex64).contiguous()
                      for t in (wavefront1, wavefront2, H_d2))
        shifts = _shifts_for(shiftsPos, self.layout, self.device)
        spec, x_fwd, sensor, amp = native_forward(w1, w2, None, h2, shifts, self.mag)
        self._saved = (spec, x_fwd, sensor, shifts)
        return amp

    def backward(self, dl_dout, wavefront1, wavefront2, H_d2, shiftsPos, observe):
        """``[dldw1, dldw2, dldH2, dldPos]`` for the last ``forward`` (the MATLAB
        class reads its buffers the same way); ``dldH2`` and ``dldPos`` are
        ``None`` (empty in MATLAB)."""
        if self._saved is None:
            raise RuntimeError("backward needs a preceding forward")
        spec, x_fwd, sensor, shifts = self._saved
        if dl_dout.shape != observe.shape or observe.shape[0] != shifts.shape[0]:
            raise ValueError("dl_dout and observe must match the last forward's output shape")
        w2 = wavefront2.to(device=self.device, dtype=torch.complex64).contiguous()
        h2 = H_d2.to(device=self.device, dtype=torch.complex64).contiguous()
        g = dl_dout.to(device=self.device, dtype=torch.float32).contiguous()
        dldw1, dldw2 = native_backward(g, spec, x_fwd, sensor, w2, None, h2, shifts, self.mag,
                                       flags=_lib.PTX_FLAG_MATLAB_NEARF)
        return dldw1, dldw2, None, None

    def delete(self) -> None:
        """Release the kept fields (the MATLAB class's ``delete``)."""
        self._saved = None

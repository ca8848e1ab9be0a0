"""Alternative losses and regularisers of the reference drivers.

* :func:`fd_loss` -- ``fd_loss(pred, target, isotropic=True)`` of
  python/train_coded.py:101-107 (selected there by ``--loss fd-isotropic`` /
  ``fd-anisotropic``): finite differences of the residual with wrap-around,
  ``sum sqrt(dx^2 + dy^2 + 1e-5)`` or ``sum |dx| + |dy|``.
* The MATLAB examples' loss functions (matlab/Examples/Ptychography/+loss_fun):
  :func:`l2_loss`, :func:`fd_loss_matlab` (``fd_loss.m``, anisotropic),
  :func:`tv_loss`, :func:`sparse_loss` and :func:`hessian_loss`, with their
  ``'sum'`` / ``'mean'`` and ``'isotropic'`` / ``'anisotropic'`` / ``'L2'``
  options.

``use_cuda_kernels=True`` on float32 / complex64 GPU tensors evaluates the loss
and its gradient in one HIP stencil kernel (``ptx_amplitude_loss`` /
``ptx_field_regularizer``); the gradient is kept for the backward, so the
autograd graph holds no intermediate differences.  Otherwise the torch ops of
the reference formulas run.  ``hessian_loss`` (second differences of an
object, used only as a regulariser) runs as torch ops.
"""

from __future__ import annotations

import torch

from . import _lib

_KIND = {
    "l2": _lib.PTX_LOSS_L2,
    "fd_iso": _lib.PTX_LOSS_FD_ISO,
    "fd_aniso": _lib.PTX_LOSS_FD_ANISO,
    "grad_l2": _lib.PTX_LOSS_GRAD_L2,
    "sparse_iso": _lib.PTX_LOSS_SPARSE_ISO,
    "sparse_aniso": _lib.PTX_LOSS_SPARSE_ANISO,
}
FD_EPS = 1e-5  # python/train_coded.py:106; +loss_fun/sparse_loss.m:9


def _diffs(d: torch.Tensor):
    dx = torch.roll(d, shifts=-1, dims=-1) - d
    dy = torch.roll(d, shifts=-1, dims=-2) - d
    return dx, dy


def _torch_penalty(kind: str, d: torch.Tensor, eps: float) -> torch.Tensor:
    if kind == "l2":
        return d.abs().square().sum()
    dx, dy = _diffs(d)
    if kind == "fd_iso":
        return torch.sqrt(dx.abs().square() + dy.abs().square() + eps).sum()
    if kind == "fd_aniso":
        return (dx.abs() + dy.abs()).sum()
    if kind == "grad_l2":
        return (dx.abs().square() + dy.abs().square()).sum()
    if kind == "sparse_iso":
        return (1 - torch.exp(-8.0 * torch.sqrt(dx.abs().square() + dy.abs().square() + eps))).sum()
    if kind == "sparse_aniso":
        return (1 - torch.exp(-8.0 * (dx.abs() + dy.abs()))).sum()
    raise ValueError(kind)


class _StencilLoss(torch.autograd.Function):
    """scale * sum of the penalty over all frames; gradient computed in the
    same kernel and saved for the backward."""

    @staticmethod
    def forward(ctx, pred, target, kind, scale, eps):  # type: ignore[override]
        cplx = pred.is_complex()
        p = pred.contiguous()
        t = target.contiguous() if target is not None else None
        h, w = p.shape[-2], p.shape[-1]
        B = p.numel() // max(h * w, 1)
        dev = p.device
        per_frame = torch.empty(B, dtype=torch.float32, device=dev)
        grad = torch.empty_like(p)
        ws = _lib.workspace(_lib.query("ptx_loss_workspace", B, h, w), dev)
        if cplx:
            _lib.call("ptx_field_regularizer", _KIND[kind], _lib.ptr(p), B, h, w, float(scale), float(eps),
                      _lib.ptr(per_frame), _lib.ptr(grad), _lib.ptr(ws), ws.numel(), _lib.stream_of(dev))
        else:
            _lib.call("ptx_amplitude_loss", _KIND[kind], _lib.ptr(p), _lib.ptr(t), B, h, w, float(scale),
                      float(eps), _lib.ptr(per_frame), _lib.ptr(grad), _lib.ptr(ws), ws.numel(), _lib.stream_of(dev))
        ctx.save_for_backward(grad)
        ctx.has_target = target is not None
        return per_frame.sum()

    @staticmethod
    def backward(ctx, g):  # type: ignore[override]
        (grad,) = ctx.saved_tensors
        gp = grad * g
        gt = -gp if ctx.has_target and ctx.needs_input_grad[1] else None
        return gp, gt, None, None, None


def _native_ok(x: torch.Tensor, use_cuda_kernels: bool) -> bool:
    return bool(use_cuda_kernels) and x.is_cuda and x.dtype in (torch.float32, torch.complex64) and x.ndim >= 2


def _stencil(kind: str, pred: torch.Tensor, target: torch.Tensor | None, scale: float, eps: float,
             use_cuda_kernels: bool) -> torch.Tensor:
    if pred.ndim < 2:
        raise ValueError("expected frames with at least two dimensions [..., h, w]")
    if target is not None:
        if target.shape != pred.shape:
            raise ValueError("pred and target must have the same shape")
        if target.device != pred.device:
            raise ValueError("pred and target must live on the same device")
    if _native_ok(pred, use_cuda_kernels) and (target is None or target.dtype == pred.dtype):
        if pred.is_complex() and target is not None:
            return _StencilLoss.apply(pred - target, None, kind, scale, eps)
        return _StencilLoss.apply(pred, target, kind, scale, eps)
    d = pred - target if target is not None else pred
    return _torch_penalty(kind, d, eps) * scale


def _reduction(kind_name: str, type_: str, numel: int) -> float:
    if type_ == "sum":
        return 1.0
    if type_ == "mean":
        return 1.0 / numel
    raise ValueError(f"{kind_name}: type must be 'sum' or 'mean'")


def fd_loss(pred: torch.Tensor, target: torch.Tensor, isotropic: bool = True,
            use_cuda_kernels: bool = False) -> torch.Tensor:
    """python/train_coded.py:101-107: with ``diff = pred - target`` and
    ``dx, dy`` its wrapped forward differences along the last and second-to-last
    axes, ``sum sqrt(dx^2 + dy^2 + 1e-5)`` (isotropic) or ``sum |dx| + |dy|``."""
    return _stencil("fd_iso" if isotropic else "fd_aniso", pred, target, 1.0, FD_EPS if isotropic else 0.0,
                    use_cuda_kernels)


def l2_loss(x1: torch.Tensor, x2: torch.Tensor, type: str = "sum", use_cuda_kernels: bool = False) -> torch.Tensor:
    """+loss_fun/l2_loss.m: ``(x1 - x2).^2`` summed or averaged."""
    return _stencil("l2", x1, x2, _reduction("l2_loss", type, x1.numel()), 0.0, use_cuda_kernels)


def fd_loss_matlab(x1: torch.Tensor, x2: torch.Tensor, type: str = "sum",
                   use_cuda_kernels: bool = False) -> torch.Tensor:
    """+loss_fun/fd_loss.m: ``abs(w([2:end,1],:,:) - w) + abs(w(:,[2:end,1],:) - w)``
    of ``w = x1 - x2``, summed or averaged (the anisotropic finite-difference loss)."""
    return _stencil("fd_aniso", x1, x2, _reduction("fd_loss", type, x1.numel()), 0.0, use_cuda_kernels)


def tv_loss(dW: torch.Tensor, type: str = "isotropic", use_cuda_kernels: bool = False) -> torch.Tensor:
    """+loss_fun/tv_loss.m: total variation of ``dW`` with wrap-around,
    ``sqrt(|Rx|^2 + |Ry|^2)`` ('isotropic', zero gradient at flat sites),
    ``|Rx| + |Ry|`` ('anisotropic') or ``|Rx|^2 + |Ry|^2`` ('L2'), summed."""
    kind = {"isotropic": "fd_iso", "anisotropic": "fd_aniso", "L2": "grad_l2"}.get(type)
    if kind is None:
        raise ValueError("tv_loss: type must be 'isotropic', 'anisotropic' or 'L2'")
    return _stencil(kind, dW, None, 1.0, 0.0, use_cuda_kernels)


def sparse_loss(dW: torch.Tensor, type: str = "isotropic", use_cuda_kernels: bool = False) -> torch.Tensor:
    """+loss_fun/sparse_loss.m: ``1 - exp(-8 sqrt(|dx|^2 + |dy|^2 + 1e-5))``
    ('isotropic'), ``1 - exp(-8 (|dx| + |dy|))`` ('anisotropic') or
    ``|dx|^2 + |dy|^2`` ('L2'), summed."""
    kind = {"isotropic": "sparse_iso", "anisotropic": "sparse_aniso", "L2": "grad_l2"}.get(type)
    if kind is None:
        raise ValueError("sparse_loss: type must be 'isotropic', 'anisotropic' or 'L2'")
    return _stencil(kind, dW, None, 1.0, FD_EPS if type == "isotropic" else 0.0, use_cuda_kernels)


def hessian_loss(dW: torch.Tensor, type: str = "isotropic") -> torch.Tensor:
    """+loss_fun/hessian_loss.m: second differences of the wrapped first
    differences, ``sqrt(|dxx|^2 + |dyy|^2 + 2|dxy|^2)`` ('isotropic'),
    ``|dxx| + |dyy| + 2|dxy|`` ('anisotropic'), or ``|dx|^2 + |dy|^2`` ('L2'),
    summed (torch ops)."""
    dx, dy = _diffs(dW)
    if type == "L2":
        return (dx.abs().square() + dy.abs().square()).sum()
    dxx = torch.roll(dx, shifts=-1, dims=-1) - dx
    dyy = torch.roll(dy, shifts=-1, dims=-2) - dy
    dxy = torch.roll(dx, shifts=-1, dims=-2) - dx
    if type == "isotropic":
        return torch.sqrt(dxx.abs().square() + dyy.abs().square() + 2 * dxy.abs().square()).sum()
    if type == "anisotropic":
        return (dxx.abs() + dyy.abs() + 2 * dxy.abs()).sum()
    raise ValueError("hessian_loss: type must be 'isotropic', 'anisotropic' or 'L2'")


def native_amplitude_loss(kind: str, pred: torch.Tensor, target: torch.Tensor, scale: float, eps: float,
                          loss_per_frame: torch.Tensor, grad: torch.Tensor, workspace: torch.Tensor) -> None:
    """Raw call of ``ptx_amplitude_loss`` into preallocated outputs (the loss
    step of a fused training step; no autograd)."""
    h, w = pred.shape[-2], pred.shape[-1]
    B = pred.numel() // (h * w)
    _lib.call("ptx_amplitude_loss", _KIND[kind], _lib.ptr(pred), _lib.ptr(target), B, h, w, float(scale), float(eps),
              _lib.ptr(loss_per_frame), _lib.ptr(grad), _lib.ptr(workspace), workspace.numel(),
              _lib.stream_of(pred.device))


__all__ = ["fd_loss", "l2_loss", "fd_loss_matlab", "tv_loss", "sparse_loss", "hessian_loss", "FD_EPS",
           "native_amplitude_loss"]

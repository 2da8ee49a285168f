"""The ``"mi355x"`` backend: torch device buffers + hand-written gfx950 HIP kernels.

Registered beside the reference's ``"numpy"`` and ``"cupy"`` names
(sgh/backends/numpy_backend.py:33-38, cupy_backend.py:37-41). ``xp`` is a thin
torch-backed namespace covering the array calls a solver needs to hold and inspect
state; all arithmetic of the step runs in ``kernels`` through the C ABI.
"""

from __future__ import annotations

import types

import numpy as np

from ._registry import register_backend
from .base import MathBackend

__all__ = ["MI355XBackend", "make_xp"]


def _torch_dtype(dtype):
    import torch
    if dtype in (None, np.float32, "float32", torch.float32):
        return torch.float32
    if dtype in (np.float64, "float64", torch.float64):
        return torch.float64
    return dtype


def make_xp(device):
    """A NumPy-like namespace over torch tensors on ``device``."""
    import torch

    def zeros(shape, dtype=np.float32):
        return torch.zeros(shape, dtype=_torch_dtype(dtype), device=device)

    def empty(shape, dtype=np.float32):
        return torch.empty(shape, dtype=_torch_dtype(dtype), device=device)

    def full(shape, value, dtype=np.float32):
        return torch.full(shape, float(value), dtype=_torch_dtype(dtype), device=device)

    def asarray(x, dtype=None):
        if isinstance(x, torch.Tensor):
            return x.to(device=device, dtype=_torch_dtype(dtype) if dtype is not None else x.dtype)
        arr = np.asarray(x)
        dt = _torch_dtype(dtype) if dtype is not None else (
            torch.float64 if arr.dtype == np.float64 else torch.float32)
        return torch.as_tensor(np.ascontiguousarray(arr), device=device).to(dt)

    def copyto(dst, src):
        dst.copy_(src)

    def vdot(a, b):
        return torch.dot(a.reshape(-1).to(torch.float64), b.reshape(-1).to(torch.float64))

    def _t(x):
        return x if isinstance(x, torch.Tensor) else torch.as_tensor(x, device=device)

    def arange(n, dtype=None):
        out = torch.arange(n, device=device)
        return out.to(_torch_dtype(dtype)) if dtype is not None else out

    def where(cond, x, y):
        c = _t(cond)
        return torch.where(c, _t(x).to(device), _t(y).to(device))

    def power(x, y):
        return torch.pow(_t(x), y)

    def gradient(f, *spacing, axis=None):
        # numpy.gradient with uniform spacings and edge_order 1 (the reference's use)
        axes = list(range(f.dim())) if axis is None else [axis]
        sp = list(spacing) if spacing else [1.0] * len(axes)
        if len(sp) == 1 and len(axes) > 1:
            sp = sp * len(axes)
        out = [torch.gradient(f, spacing=float(h), dim=ax, edge_order=1)[0]
               for ax, h in zip(axes, sp)]
        return out[0] if axis is not None or len(out) == 1 else out

    def tensordot(a, b, axes=2):
        return torch.tensordot(a, b, dims=axes)

    def einsum(spec, *ops, optimize=None):
        return torch.einsum(spec, *ops)

    ns = types.SimpleNamespace(
        zeros=zeros, empty=empty, full=full, asarray=asarray, copyto=copyto, vdot=vdot,
        zeros_like=torch.zeros_like, empty_like=torch.empty_like,
        sum=torch.sum, max=torch.max, abs=torch.abs, sqrt=torch.sqrt,
        maximum=torch.maximum, multiply=torch.mul, subtract=torch.sub,
        exp=lambda x: torch.exp(_t(x)), cos=lambda x: torch.cos(_t(x)),
        arange=arange, where=where, power=power, gradient=gradient, tensordot=tensordot,
        einsum=einsum, pi=np.pi,
        float32=np.float32, float64=np.float64, __name__="mi355x_xp",
    )
    return ns


@register_backend("mi355x")
class MI355XBackend(MathBackend):
    """GPU backend for MI355X: ``kernels`` is ``physics.mi355x_kernels``."""

    def __init__(self, device: str = "cuda:0"):
        import torch

        from .. import _lib
        from ..physics import mi355x_kernels
        if not torch.cuda.is_available():
            raise ImportError("the mi355x backend needs a ROCm GPU (torch.cuda is unavailable)")
        _lib.lib.load()
        super().__init__(name="mi355x", xp=make_xp(device), kernels=mi355x_kernels)
        self.device = device

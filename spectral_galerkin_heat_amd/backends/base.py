"""``MathBackend``: an array module plus a kernel module (sgh/backends/base.py:28-81)."""

from __future__ import annotations

import numpy as _np

__all__ = ["MathBackend", "to_host"]


def to_host(arr):
    """Return *arr* as a NumPy array on the host (torch device tensors are copied back)."""
    if hasattr(arr, "detach") and hasattr(arr, "cpu"):
        return arr.detach().cpu().numpy()
    get = getattr(arr, "get", None)
    return get() if callable(get) else _np.asarray(arr)


class MathBackend:
    """Hold the array module ``xp`` and the physics ``kernels`` for one execution target."""

    def __init__(self, name: str, xp, kernels):
        self.name = name
        self.xp = xp
        self.kernels = kernels

    def to_numpy(self, arr):
        return to_host(arr)

    def __repr__(self) -> str:
        return f"MathBackend(name={self.name!r})"

"""Error norms of two temperature fields on a structured grid, as the reference defines them.

Restates ``compute_L2_structured`` and ``_trapezoidal_weights``
(sgh/io_utils/compute_L2_error.py:436-442, 491-530): composite trapezoid weights per axis
(half a spacing at each end of a non-uniform coordinate list), their tensor product
``W[z, y, x]``, ``L2_abs = sqrt(Σ W·(T_a − T_b)²)``, ``L2_rel = L2_abs / sqrt(Σ W·T_a²)``,
``Linf = max |T_a − T_b|``, and the same error on non-finite values. Fields are indexed
``[z, y, x]``. NumPy arrays are summed in float64 on the host; torch tensors (e.g. two
node-centred reconstructions still on the device) are summed in float64 where they lie.
"""

from __future__ import annotations

import numpy as np

__all__ = ["trapezoidal_weights", "compute_L2_structured"]


def trapezoidal_weights(coords) -> np.ndarray:
    """Composite trapezoid quadrature weights of a (possibly non-uniform) coordinate list."""
    c = np.asarray(coords, dtype=np.float64)
    w = np.zeros_like(c)
    if c.size < 2:
        return w
    h = np.diff(c)
    w[:-1] += 0.5 * h
    w[1:] += 0.5 * h
    return w


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def compute_L2_structured(T_a, T_b, x, y, z) -> dict:
    """Absolute / relative L2 and L∞ of ``T_a − T_b`` (both ``[z, y, x]``) with trapezoid
    weights on the node coordinates ``x``, ``y``, ``z``; ``T_a`` is the reference field."""
    wx, wy, wz = (trapezoidal_weights(c) for c in (x, y, z))
    shape = (len(wz), len(wy), len(wx))
    if tuple(T_a.shape) != shape or tuple(T_b.shape) != shape:
        raise ValueError(f"fields of shape {tuple(T_a.shape)} / {tuple(T_b.shape)} do not match "
                         f"the coordinates' (z, y, x) = {shape}")
    if _is_torch(T_a) or _is_torch(T_b):
        import torch
        dev = T_a.device if _is_torch(T_a) else T_b.device
        a = torch.as_tensor(T_a, device=dev).double()
        b = torch.as_tensor(T_b, device=dev).double()
        W = (torch.as_tensor(wz, device=dev)[:, None, None]
             * torch.as_tensor(wy, device=dev)[None, :, None]
             * torch.as_tensor(wx, device=dev)[None, None, :])
        err = a - b
        if not bool(torch.isfinite(err).all()):
            raise ValueError("Error fields contain NaN values, indicating a severe issue "
                             "(e.g. out of bounds interpolation).")
        volume = float(W.sum())
        L2_sq = float((W * err * err).sum())
        ref_sq = float((W * a * a).sum())
        Linf = float(err.abs().max())
    else:
        a = np.asarray(T_a, dtype=np.float64)
        b = np.asarray(T_b, dtype=np.float64)
        W = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        err = a - b
        if not np.isfinite(err).all():
            raise ValueError("Error fields contain NaN values, indicating a severe issue "
                             "(e.g. out of bounds interpolation).")
        volume = float(W.sum())
        L2_sq = float(np.sum(W * err ** 2))
        ref_sq = float(np.sum(W * a ** 2))
        Linf = float(np.max(np.abs(err)))
    L2_abs = L2_sq ** 0.5
    n_total = int(np.prod(shape))
    return {"L2_abs": L2_abs,
            "L2_rel": L2_abs / ref_sq ** 0.5 if ref_sq > 0 else float("inf"),
            "Linf": Linf, "volume": volume, "n_valid": n_total, "n_total": n_total,
            "pct_valid": 100.0, "n_nan": 0}

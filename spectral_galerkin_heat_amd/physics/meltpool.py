"""Melt-pool extent of a cell-centred temperature field.

The reference has no melt-pool function; the extent used to compare precision forms is
the bounding box of the cells with ``T >= T_melt`` on the cell-centred grid
(``reconstruct_volume``, sgh/physics/spectral_ops.py:83-91): length along x and width along
y count the melted cells between the outermost ones, depth counts the layers from the top
face (z = Lz, the last z index) down to the deepest melted cell.
"""

from __future__ import annotations

import numpy as np
import torch

__all__ = ["cell_temperature", "melt_pool_extent"]


def cell_temperature(a, geom):
    """Cell-centred temperature of the modal state ``a`` (nz, ny, nx) on the device:
    three MFMA axis contractions with the cosine bases (reconstruct_volume)."""
    from . import mi355x_kernels as K
    g = K.SpectralGrid(geom, np.float64 if a.dtype == torch.float64 else np.float32)
    return K._apply_axes(a.contiguous(), (g.BxT, g.ByT, g.BzT))


def melt_pool_extent(T, d, T_melt=1723.0):
    """Return ``(length, width, depth)`` in metres; zeros when nothing has melted.

    ``T`` is a ``(nz, ny, nx)`` tensor (any device), ``d = (dx, dy, dz)``.
    """
    mask = torch.as_tensor(T) >= T_melt
    if not bool(mask.any()):
        return 0.0, 0.0, 0.0
    nz = mask.shape[0]
    idx = torch.nonzero(mask)
    kmin = int(idx[:, 0].min())
    jmin, jmax = int(idx[:, 1].min()), int(idx[:, 1].max())
    imin, imax = int(idx[:, 2].min()), int(idx[:, 2].max())
    dx, dy, dz = (float(v) for v in d)
    return (imax - imin + 1) * dx, (jmax - jmin + 1) * dy, (nz - kmin) * dz

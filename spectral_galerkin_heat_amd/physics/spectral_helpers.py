"""Output-side helpers of the MI355X backend (sgh/physics/spectral_helpers.py).

``reconstruct_temperature_DCT`` evaluates the modal state at the node-centred grid
x_j = j·Lx/nx (j = 0..nx), y, z likewise, the field the reference's HDF5/XDMF writers
store. The reference pads the modes to n+1, halves the interior ones and runs an
unnormalised DCT-I (spectral_helpers.py:147-207); here each axis is one MFMA contraction
with the (n+1) × n node basis N[j, k] = C_k·cos(πkj/n), on the device.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .._lib import call, dtype_code, ptr, stream_ptr

__all__ = ["_calculate_subgrid_indices", "node_basis", "reconstruct_temperature_DCT",
           "reconstruct_temperature_volume", "reconstruct_temperature_volume_at_points"]

_BASES: dict = {}


def _calculate_subgrid_indices(pos, dx, n_total_fine, n_box):
    """Window of ``n_box`` fine cells centred on ``pos`` (spectral_helpers.py:49-83).

    Returns ``(start, end, offset)``: the centre cell is ``round(pos/dx)`` (Python's
    round, half to even) clamped to the fine grid, the start is clamped so the window
    stays inside the grid when it fits (start 0 otherwise), and ``offset`` is the centre
    cell's index inside the window, clamped to it.
    """
    centre = round(max(0.0, min(float(pos) / float(dx), n_total_fine - 1)))
    start = max(0, min(centre - n_box // 2, max(0, n_total_fine - n_box)))
    return start, start + n_box, max(0, min(centre - start, n_box - 1))


def node_basis(n: int, L: float) -> np.ndarray:
    """N[j, k] = C_k·cos(πkj/n) for j = 0..n, k = 0..n-1 (FP64, host)."""
    j = np.arange(n + 1)[:, None]
    k = np.arange(n)[None, :]
    ang = np.pi * ((k * j) % (2 * n)) / n
    C = np.full(n, math.sqrt(2.0 / L))
    C[0] = math.sqrt(1.0 / L)
    return np.cos(ang) * C[None, :]


def _basis(n, L, tdtype, device):
    key = (n, float(L), tdtype, str(device))
    B = _BASES.get(key)
    if B is None:
        B = torch.as_tensor(node_basis(n, L)).to(device=device, dtype=tdtype).contiguous()
        _BASES[key] = B
    return B


def reconstruct_temperature_DCT(a, SsState):
    """Node-centred temperature, shape (N_x+1, N_y+1, N_z+1), from modes ``a`` (nz, ny, nx).

    ``SsState`` supplies the domain (``grid.nx/ny/nz`` and the extents through
    ``grid.dx·nx`` etc.); the result stays on the device in the modes' dtype.
    """
    a = a.contiguous()
    nz, ny, nx = a.shape
    g = SsState.grid
    Ls = (g.dx * nx, g.dy * ny, g.dz * nz)
    code = dtype_code(a.dtype)
    st = stream_ptr()
    Bx, By, Bz = (_basis(n, L, a.dtype, a.device) for n, L in zip((nx, ny, nz), Ls))
    t1 = torch.empty((nz, ny, nx + 1), dtype=a.dtype, device=a.device)
    call("sgh_axis_contract_rect", code, 0, nz * ny, nx, nx + 1, 1, ptr(Bx), ptr(a), 1.0,
         ptr(t1), st)
    t2 = torch.empty((nz, ny + 1, nx + 1), dtype=a.dtype, device=a.device)
    call("sgh_axis_contract_rect", code, 0, nz, ny, ny + 1, nx + 1, ptr(By), ptr(t1), 1.0,
         ptr(t2), st)
    t3 = torch.empty((nz + 1, ny + 1, nx + 1), dtype=a.dtype, device=a.device)
    call("sgh_axis_contract_rect", code, 0, 1, nz, nz + 1, (ny + 1) * (nx + 1), ptr(Bz),
         ptr(t2), 1.0, ptr(t3), st)
    return t3.permute(2, 1, 0).contiguous()


def reconstruct_temperature_volume(a, SsState):
    """The node-centred field of ``reconstruct_temperature_DCT`` (spectral_helpers.py:108-144
    evaluates the same series as a direct tensor product; here both are the same three
    device contractions)."""
    return reconstruct_temperature_DCT(a, SsState)


def reconstruct_temperature_volume_at_points(a, num, geom, SsState, coords):
    """T at arbitrary points (spectral_helpers.py:209-255).

    ``coords`` is an (N, 3) array of (x, y, z) in metres, clipped to the domain. The bases
    at the points are built on the host in FP64 (as the reference builds them), the x
    contraction is one MFMA launch (``sgh_axis_contract_rect``) and the y, z sums one
    thread per point (``sgh_points_yz``); the result is an (N,) device tensor in the
    modes' dtype.
    """
    coords = np.asarray(coords)
    if coords.ndim != 2 or coords.shape[1] != 3:
        raise ValueError("coords must be of shape (N, 3)")
    a = a.contiguous()
    nz, ny, nx = a.shape
    Ls = (float(geom.size.x), float(geom.size.y), float(geom.size.z))
    P = coords.shape[0]
    if P == 0:
        return torch.empty(0, dtype=a.dtype, device=a.device)
    vals = [np.clip(coords[:, k].astype(np.float64), 0.0, Ls[k]) for k in range(3)]

    def basis(n, L, v):
        C = np.full(n, math.sqrt(2.0 / L))
        C[0] = math.sqrt(1.0 / L)
        return C[:, None] * np.cos(np.pi * np.arange(n)[:, None] * v[None, :] / L)

    dev = lambda m: torch.as_tensor(np.ascontiguousarray(m)).to(a.device, a.dtype)  # noqa: E731
    BxT = dev(basis(nx, Ls[0], vals[0]).T)          # (P, nx): the x contraction's matrix
    By, Bz = dev(basis(ny, Ls[1], vals[1])), dev(basis(nz, Ls[2], vals[2]))
    code, st = dtype_code(a.dtype), stream_ptr()
    U = torch.empty((nz, ny, P), dtype=a.dtype, device=a.device)
    call("sgh_axis_contract_rect", code, 0, nz * ny, nx, P, 1, ptr(BxT), ptr(a), 1.0, ptr(U), st)
    out = torch.empty(P, dtype=a.dtype, device=a.device)
    call("sgh_points_yz", code, nz, ny, P, ptr(U), ptr(By), ptr(Bz), ptr(out), st)
    return out

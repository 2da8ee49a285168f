"""MI355X kernel module: the reference kernel-module API on torch device buffers.

Same ``__all__`` as the reference's kernel modules (``sgh/physics/spectral_cpu_kernels.py:37-48``,
``spectral_gpu_kernels.py:36-47``) plus the extra module-level names the solver reaches
(``DCT_II``, ``IDCT_II``, ``shift_flux``, ``compute_evaporation_flux``,
``reconstruct_volume``, ``project_volume``, ``assemble_property_correction``, ...).
Every function calls a gfx950 HIP kernel through the C ABI (``_lib``); none falls back
to host arithmetic. ``picard_step`` is the fused per-step driver the ``SpectralSolver``
uses on this backend.

Both latent-heat modes: grid mode (latent source on the coarse grid, projected by the
3-D transforms) and box mode (a laser-following fine sub-box whose bases stay resident
in HBM; each step's window is a column offset read in place by the strided MFMA
contraction).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .. import _lib
from .._lib import call, dtype_code, ptr, stream_ptr

__all__ = [
    "FineMeshState",
    "SpectralGrid",
    "SpectralSolverState",
    "add_bottom_surface_source",
    "add_source_term_modes",
    "compute_latent_heat_source",
    "compute_latent_heat_source_grid",
    "reconstruct_bottom_temperature",
    "reconstruct_surface_temperature",
    "update_modes_etd1",
]

_DEVICE = "cuda"


def _tdtype(dtype):
    return torch.float64 if np.dtype(dtype) == np.float64 else torch.float32


def _dev(arr, dtype):
    return torch.as_tensor(np.ascontiguousarray(arr, dtype=np.float64)).to(
        device=_DEVICE, dtype=_tdtype(dtype)).contiguous()


def cosine_basis(n: int, L: float) -> np.ndarray:
    """B[k, i] = C_k·cos(πk(i+½)/n) with C = √(2/L), √(1/L) for k = 0 (FP64, host).

    The surface reconstruction is Bᵧᵀ·A·Bₓ and the projection dx·dy·Bᵧ·q·Bₓᵀ, which equal
    the reference's recon_scale·IDCT_II and dct_scale·DCT_II (spectral_state.py:115-116).
    """
    k = np.arange(n)[:, None]
    i = np.arange(n)[None, :]
    ang = np.pi * ((k * (2 * i + 1)) % (4 * n)) / (2.0 * n)
    C = np.full(n, math.sqrt(2.0 / L))
    C[0] = math.sqrt(1.0 / L)
    return C[:, None] * np.cos(ang)


def dct_ortho_matrix(n: int) -> np.ndarray:
    """Ortho DCT-II matrix M[k, j] = s_k·cos(πk(2j+1)/(2n)), s_0 = √(1/n), s_k = √(2/n)."""
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    ang = np.pi * ((k * (2 * j + 1)) % (4 * n)) / (2.0 * n)
    s = np.full(n, math.sqrt(2.0 / n))
    s[0] = math.sqrt(1.0 / n)
    return s[:, None] * np.cos(ang)


_DCT_CACHE: dict = {}


def _dct_mats(n, tdtype):
    key = (n, tdtype)
    m = _DCT_CACHE.get(key)
    if m is None:
        M = dct_ortho_matrix(n)
        m = (torch.as_tensor(M).to(_DEVICE, tdtype).contiguous(),
             torch.as_tensor(np.ascontiguousarray(M.T)).to(_DEVICE, tdtype).contiguous())
        _DCT_CACHE[key] = m
    return m


TRANSFORM_VARIANTS = {"mfma": 0, "bf16x3": 1, "valu": 2}


def transform_code(name, dtype):
    """C-ABI variant code of a 3-D transform arithmetic name (SGH_XFORM_*)."""
    key = str(name).lower()
    if key not in TRANSFORM_VARIANTS:
        raise ValueError(f"unknown transform {name!r}; choose one of {sorted(TRANSFORM_VARIANTS)}")
    if key == "bf16x3" and _tdtype(dtype) != torch.float32:
        raise ValueError("the split-BF16 transform runs on float32 data (simulation.dtype float32)")
    return TRANSFORM_VARIANTS[key]


def _apply_axes(x, mats, scale=1.0, xform=0):
    """Contract every axis of a 3-D tensor with its matrix (mats in x, y, z order).

    Three MFMA GEMM launches (``sgh_axis_contract``): x as Out[zy, k] = In[zy, j]·M[k, j],
    y batched over z planes, z as one GEMM over all (y, x) columns.
    """
    nz, ny, nx = x.shape
    code = dtype_code(x.dtype)
    st = stream_ptr()
    w1 = torch.empty_like(x)
    w2 = torch.empty_like(x)
    call("sgh_axis_contract", code, xform, nz * ny, nx, 1, ptr(mats[0]), ptr(x), 1.0, ptr(w1), st)
    call("sgh_axis_contract", code, xform, nz, ny, nx, ptr(mats[1]), ptr(w1), 1.0, ptr(w2), st)
    call("sgh_axis_contract", code, xform, 1, nz, ny * nx, ptr(mats[2]), ptr(w2), float(scale),
         ptr(w1), st)
    return w1


def _sep2d(X, L, R, scale=1.0, nparts=1, out=None):
    ny, nx = L.shape[0], R.shape[0]
    if out is None:
        out = torch.empty((ny, nx), dtype=X.dtype, device=X.device)
    call("sgh_sep2d", dtype_code(X.dtype), 1, ny, nx, ptr(L), ptr(R), ptr(X), int(nparts),
         float(scale), ptr(out), stream_ptr())
    return out


def DCT_II(q):
    """Ortho DCT-II over every axis (spectral_cpu_kernels.py:374-381).

    2-D: one LDS GEMM-pair launch; 3-D: three MFMA axis contractions in the data type."""
    q = q.contiguous()
    td = q.dtype
    if q.dim() == 2:
        My, _ = _dct_mats(q.shape[0], td)
        _, MxT = _dct_mats(q.shape[1], td)
        return _sep2d(q, My, MxT)
    if q.dim() == 3:
        mats = [_dct_mats(n, td)[0] for n in (q.shape[2], q.shape[1], q.shape[0])]
        return _apply_axes(q, mats)
    raise ValueError("DCT_II supports 2-D and 3-D fields")


def IDCT_II(a):
    """Ortho DCT-III over every axis, the inverse of DCT_II (spectral_cpu_kernels.py:383-389).

    Same launch structure as ``DCT_II``."""
    a = a.contiguous()
    td = a.dtype
    if a.dim() == 2:
        _, MyT = _dct_mats(a.shape[0], td)
        Mx, _ = _dct_mats(a.shape[1], td)
        return _sep2d(a, MyT, Mx)
    if a.dim() == 3:
        mats = [_dct_mats(n, td)[1] for n in (a.shape[2], a.shape[1], a.shape[0])]
        return _apply_axes(a, mats)
    raise ValueError("IDCT_II supports 2-D and 3-D fields")


# ---------------------------------------------------------------------------------
# State
# ---------------------------------------------------------------------------------


class SpectralGrid:
    """Grid constants and cosine bases on the device (spectral_state.py:71-143)."""

    def __init__(self, geom, dtype=np.float32):
        self.dtype = dtype
        td = _tdtype(dtype)
        nx, ny, nz = (int(v) for v in geom.n)
        Lx, Ly, Lz = (float(v) for v in geom.size)
        dx, dy, dz = (float(v) for v in geom.d)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.x, self.y, self.z = [_dev((np.arange(n) + 0.5) * d, dtype)
                                  for n, d in ((nx, dx), (ny, dy), (nz, dz))]
        Cs = []
        for n, L in ((nx, Lx), (ny, Ly), (nz, Lz)):
            c = np.full(n, math.sqrt(2.0 / L))
            c[0] = math.sqrt(1.0 / L)
            Cs.append(c)
        self.C = tuple(_dev(c, dtype) for c in Cs)
        self.dct_scale = dtype((dx * dy) * np.sqrt((nx * ny) / (Lx * Ly)))
        self.recon_scale = dtype(np.sqrt(nx * ny) / np.sqrt(Lx * Ly))
        sign = np.power(-1.0, np.arange(nz))
        self.w_top = _dev(Cs[2] * sign, dtype)
        self.w_bot = _dev(Cs[2], dtype)
        self.Cp32_broadcast = self.w_top.view(nz, 1, 1)
        self.Cp32_broadcast_bottom = self.w_bot.view(nz, 1, 1)
        self.sqrt_dV = dtype(np.sqrt(dx * dy * dz))
        self.dx, self.dy, self.dz = dx, dy, dz
        self.kx = _dev(np.pi * np.arange(nx) / Lx, dtype)
        self.ky = _dev(np.pi * np.arange(ny) / Ly, dtype)
        self.kz = _dev(np.pi * np.arange(nz) / Lz, dtype)
        B = [cosine_basis(n, L) for n, L in ((nx, Lx), (ny, Ly), (nz, Lz))]
        self.Bx, self.By, self.Bz = (_dev(b, dtype) for b in B)
        self.BxT, self.ByT, self.BzT = (_dev(np.ascontiguousarray(b.T), dtype) for b in B)
        self.dV = dx * dy * dz
        self.B_recon = None
        self.coords_rec = None


def _cosine_at(n_modes, L, coords):
    """B[k, i] = C_k·cos(π·k·coords_i/L) (FP64, host)."""
    C = np.full(n_modes, math.sqrt(2.0 / L))
    C[0] = math.sqrt(1.0 / L)
    return C[:, None] * np.cos(np.pi * np.arange(n_modes)[:, None] * coords[None, :] / L)


class FineMeshState:
    """Laser-following fine sub-box (spectral_state.py:166-239), resident on the device.

    Fine spacing ``d/refinement`` per axis. x and y: the bases over the whole fine grid
    (``modes × n_total``, and their transposes) stay in HBM; ``update`` only moves the
    window start, which the contractions read in place at row stride ``n_total``. z: the
    top ``box_size.z`` of the domain (cell centres shifted by ``Lz − Lz_box``), fixed.
    ``T_prev`` / ``Q_prev`` are the box history ``(nz_box, ny_box, nx_box)``.
    """

    def __init__(self, geom, grid, fine, dtype=np.float32):
        from .spectral_helpers import _calculate_subgrid_indices
        self._window = _calculate_subgrid_indices
        self.dtype = dtype
        td = _tdtype(dtype)
        r = int(fine.refinement)
        if r < 1:
            raise ValueError("fine_mesh.refinement must be >= 1")
        self.refinement = r
        dx, dy, dz = (float(v) for v in geom.d)
        Lx, Ly, Lz = (float(v) for v in geom.size)
        Lbx, Lby, Lbz = (float(v) for v in fine.box_size)
        if not 0.0 < Lbz <= Lz * (1.0 + 1e-12):
            raise ValueError("fine_mesh.box_size z must lie in (0, Lz]")
        self.dx_fine, self.dy_fine, self.dz_fine = dx / r, dy / r, dz / r
        d_fine = (self.dx_fine, self.dy_fine, self.dz_fine)
        self.n_fine_totals = [int(np.ceil(L / d)) for L, d in zip((Lx, Ly, Lbz), d_fine)]
        coords = [(np.arange(n) + 0.5) * d for n, d in zip(self.n_fine_totals, d_fine)]
        self.coords_fine = coords
        z_global = (Lz - Lbz) + coords[2]
        nx, ny, nz = (int(v) for v in geom.n)
        self.nx_box = min(int(np.ceil(Lbx / self.dx_fine)), self.n_fine_totals[0])
        self.ny_box = min(int(np.ceil(Lby / self.dy_fine)), self.n_fine_totals[1])
        self.nz_box = min(int(np.ceil(Lbz / self.dz_fine)), self.n_fine_totals[2])
        Bx = _cosine_at(nx, Lx, coords[0])
        By = _cosine_at(ny, Ly, coords[1])
        Bz = _cosine_at(nz, Lz, z_global)[:, :self.nz_box]
        self.Bx_full, self.By_full, self.Bz = (_dev(b, dtype) for b in (Bx, By, Bz))
        self.BxT_full, self.ByT_full, self.BzT = (_dev(np.ascontiguousarray(b.T), dtype)
                                                  for b in (Bx, By, Bz))
        self.n_modes = (nx, ny, nz)
        self.shape = (self.nz_box, self.ny_box, self.nx_box)
        self.dV_fine = self.dx_fine * self.dy_fine * self.dz_fine
        self.x_start = self.y_start = 0
        dev = dict(dtype=td, device=_DEVICE)
        # intermediates: (nz_box, ny, nx) then (nz_box, ny_box, nx) in both directions
        work = self.nz_box * max(ny, self.ny_box) * nx
        self._wbuf = torch.empty(2 * work, **dev)
        self._w = (self._wbuf[:work], self._wbuf[work:])
        self._T = torch.empty(self.shape, **dev)        # history, and its shift target
        self._T_alt = torch.empty(self.shape, **dev)
        self._Q = torch.empty(self.shape, **dev)
        self._Q_alt = torch.empty(self.shape, **dev)
        self.T_box = torch.empty(self.shape, **dev)     # per-iteration reconstruction
        self.Q_modes = torch.empty((nz, ny, nx), **dev)
        self.peaks = torch.zeros(4 * (self.nz_box + 1), dtype=torch.float64, device=_DEVICE)
        self.T_prev = None
        self.Q_prev = None

    def update(self, laser_state):
        """Move the x/y windows to the laser position (spectral_state.py:227-235)."""
        self.x_start = self._window(laser_state.x, self.dx_fine, self.n_fine_totals[0],
                                    self.nx_box)[0]
        self.y_start = self._window(laser_state.y, self.dy_fine, self.n_fine_totals[1],
                                    self.ny_box)[0]

    @property
    def B_fine(self):
        """The active window bases (x, y, z) as views, each ``modes × n_box``."""
        return [self.Bx_full[:, self.x_start:self.x_start + self.nx_box],
                self.By_full[:, self.y_start:self.y_start + self.ny_box], self.Bz]

    def _col(self, t, col):
        return t.data_ptr() + col * t.element_size()


def _reconstruct_temperature_box(a, SsState, out=None):
    """Temperature on the fine box cells, Σ a·B_z B_y B_x (spectral_ops.py:173-179).

    Three strided MFMA contractions, smallest intermediates first: z over all (y, x)
    mode columns, then y (the row block of B_yᵀ at the window start), then x likewise."""
    fm = SsState.fine_mesh
    if fm is None:
        raise RuntimeError("Fine mesh not initialized.")
    a = a.contiguous()
    nz, ny, nx = a.shape
    nzb, nyb, nxb = fm.shape
    code, xf, st = dtype_code(a.dtype), SsState.xform, stream_ptr()
    w1, w2 = fm._w
    if out is None:
        out = torch.empty(fm.shape, dtype=a.dtype, device=a.device)
    call("sgh_axis_contract_strided", code, xf, 1, nz, nzb, ny * nx, ptr(fm.BzT), nz, ptr(a),
         1.0, ptr(w1), st)
    call("sgh_axis_contract_strided", code, xf, nzb, ny, nyb, nx,
         fm._col(fm.ByT_full, fm.y_start * ny), ny, ptr(w1), 1.0, ptr(w2), st)
    call("sgh_axis_contract_strided", code, xf, nzb * nyb, nx, nxb, 1,
         fm._col(fm.BxT_full, fm.x_start * nx), nx, ptr(w2), 1.0, ptr(out), st)
    return out


def project_box_to_modes(field_box, SsState, out=None):
    """Modes of a box field, dV_fine·Σ f·B_z B_y B_x (spectral_ops.py:53-74).

    x: window columns of Bₓ read in place at row stride n_total; y likewise; z last, with
    dV_fine folded into its epilogue."""
    fm = SsState.fine_mesh
    if fm is None:
        raise RuntimeError("Fine mesh not initialized.")
    f = field_box.contiguous()
    if tuple(f.shape) != fm.shape:
        raise ValueError(f"box field has shape {tuple(f.shape)}, expected {fm.shape}")
    nzb, nyb, nxb = fm.shape
    nx, ny, nz = fm.n_modes
    code, xf, st = dtype_code(f.dtype), SsState.xform, stream_ptr()
    w1, w2 = fm._w
    if out is None:
        out = torch.empty((nz, ny, nx), dtype=f.dtype, device=f.device)
    call("sgh_axis_contract_strided", code, xf, nzb * nyb, nxb, nx, 1,
         fm._col(fm.Bx_full, fm.x_start), fm.n_fine_totals[0], ptr(f), 1.0, ptr(w1), st)
    call("sgh_axis_contract_strided", code, xf, nzb, nyb, ny, nx,
         fm._col(fm.By_full, fm.y_start), fm.n_fine_totals[1], ptr(w1), 1.0, ptr(w2), st)
    call("sgh_axis_contract_strided", code, xf, 1, nzb, nz, ny * nx, ptr(fm.Bz), nzb,
         ptr(w2), float(fm.dV_fine), ptr(out), st)
    return out


class SolverBuffers:
    """Reusable device buffers (spectral_state.py:242-275) plus the fused-step scratch."""

    def __init__(self, num, dtype, fine_mesh=None):
        td = _tdtype(dtype)
        nx, ny, nz = num.nx, num.ny, num.nz
        shp, srf = (nz, ny, nx), (ny, nx)
        dev = dict(dtype=td, device=_DEVICE)
        self.a_temp = torch.empty(shp, **dev)
        self.a_alt = torch.empty(shp, **dev)
        self.q_evap_old = torch.zeros(srf, **dev)
        self.q_evap_buffer = torch.zeros(srf, **dev)
        # box mode: the latent source lives on the fine box (spectral_state.py:259-266)
        self.Q_latent_buffer = torch.zeros(fine_mesh.shape if fine_mesh is not None else shp,
                                           **dev)
        self.n_elements = dtype(nx * ny * nz)
        self.q_las = torch.zeros(srf, **dev)
        self.q_tmp = torch.zeros(srf, **dev)
        self.T_s = torch.zeros(srf, **dev)
        self.S_top = torch.zeros(srf, **dev)
        self.S_tmp = torch.zeros(srf, **dev)
        self.S_bot = torch.zeros(srf, **dev)
        self.S_bot_raw = torch.zeros(srf, **dev)
        self.S_zero = torch.zeros(srf, **dev)
        self.nchunks = max(1, min(nz, 8))
        self.zc = torch.zeros((self.nchunks, ny, nx), **dev)
        n_pairs = _lib.lib.sgh_picard_partials_count(nx, ny, nz, self.nchunks)
        self.n_pairs = n_pairs
        self.partials = torch.zeros(2 * n_pairs, dtype=torch.float64, device=_DEVICE)
        self.norms = torch.zeros(3, dtype=torch.float64, device=_DEVICE)
        self.scal = torch.zeros(4, dtype=torch.float64, device=_DEVICE)
        self.norms_host = torch.zeros(3, dtype=torch.float64).pin_memory()


class SpectralSolverState:
    """Device state: grid, buffers, modes ``a`` and the E/KK propagators (spectral_state.py:315-365)."""

    def __init__(self, phys, geom, num, fine):
        self.dtype = getattr(num, "dtype", np.float32)
        self.code = dtype_code(self.dtype)
        self.xp = None
        self.grid = SpectralGrid(geom, self.dtype)
        self.fine_mesh = (FineMeshState(geom, self.grid, fine, self.dtype)
                          if fine is not None else None)
        self.buffers = SolverBuffers(num, self.dtype, self.fine_mesh)
        self.K, self.KK = precompute_K_KK(phys, num, geom, self.dtype)
        self.xform = transform_code(getattr(num, "transform", "mfma"), self.dtype)
        self.a = None
        self.hooks = None


def precompute_K_KK(phys, num, geom, dtype=np.float32):
    """E and KK tables on the device in FP64 math (spectral_state.py:374-421)."""
    nx, ny, nz = (int(v) for v in geom.n)
    Lx, Ly, Lz = (float(v) for v in geom.size)
    td = _tdtype(dtype)
    K = torch.empty((nz, ny, nx), dtype=td, device=_DEVICE)
    KK = torch.empty_like(K)
    alpha = float(phys.k) / (float(phys.rho) * float(phys.Cp))
    call("sgh_etd_tables", dtype_code(dtype), nx, ny, nz, Lx, Ly, Lz, alpha,
         float(phys.rho) * float(phys.Cp), float(num.dt), ptr(K), ptr(KK), stream_ptr())
    return K, KK


# ---------------------------------------------------------------------------------
# Fine-grained kernels (reference names)
# ---------------------------------------------------------------------------------


def _vec(cp):
    return cp.reshape(-1).contiguous()


def update_modes_etd1(aK, KK, Cp_broadcast, B_scaled, a_temp_out):
    """a_out = aK + KK·Cp[p]·B (spectral_cpu_kernels.py:83-91)."""
    nz, ny, nx = aK.shape
    call("sgh_etd1_update", dtype_code(aK.dtype), nx, ny, nz, ptr(aK), ptr(KK),
         ptr(_vec(Cp_broadcast)), ptr(B_scaled.contiguous()), None, None, None, ptr(a_temp_out),
         stream_ptr())


def add_source_term_modes(a_temp, KK, Q_modes):
    """a += KK·Q (spectral_cpu_kernels.py:93-103)."""
    call("sgh_add_volume_source", dtype_code(a_temp.dtype), a_temp.numel(), ptr(a_temp), ptr(KK),
         ptr(Q_modes.contiguous()), stream_ptr())


def add_bottom_surface_source(a_temp, KK, Cp_broadcast_bottom, B_scaled):
    """a += KK·Cb[p]·B (spectral_cpu_kernels.py:105-113)."""
    nz, ny, nx = a_temp.shape
    call("sgh_add_surface_source", dtype_code(a_temp.dtype), nx, ny, nz, ptr(a_temp), ptr(KK),
         ptr(_vec(Cp_broadcast_bottom)), ptr(B_scaled.contiguous()), stream_ptr())


def _recon_face(a, w, SsState, out=None):
    g, b = SsState.grid, SsState.buffers
    nz, ny, nx = a.shape
    code = dtype_code(a.dtype)
    call("sgh_zcontract", code, nx, ny, nz, ptr(a), ptr(w), b.nchunks, ptr(b.zc), stream_ptr())
    call("sgh_sum_parts", code, nx * ny, b.nchunks, ptr(b.zc), stream_ptr())
    return _sep2d(b.zc, g.ByT, g.Bx, 1.0, 1, out=out)


def reconstruct_surface_temperature(a, SsState):
    """Top-surface temperature at z = Lz (spectral_ops.py:210-219)."""
    return _recon_face(a.contiguous(), SsState.grid.w_top, SsState)


def reconstruct_bottom_temperature(a, SsState):
    """Bottom-surface temperature at z = 0 (spectral_ops.py:222-231)."""
    return _recon_face(a.contiguous(), SsState.grid.w_bot, SsState)


def compute_evaporation_flux(T_surface, q_out, P0, T_boil, DeltaH_LV, R_v, T_liquidus):
    """Evaporation flux into ``q_out`` (spectral_cpu_kernels.py:144-162)."""
    scratch = torch.empty_like(q_out)
    call("sgh_surface_forcing", dtype_code(q_out.dtype), q_out.numel(), ptr(T_surface.contiguous()),
         None, float(P0), float(T_boil), float(DeltaH_LV), float(R_v), float(T_liquidus), 0.0,
         0.0, None, 0.0, ptr(q_out), ptr(scratch), stream_ptr())


def reconstruct_volume(a, SsState):
    """T = Σ a·B_z B_y B_x on the cell-centred grid = IDCT(a)/√dV (spectral_ops.py:83-91)."""
    g = SsState.grid
    return _apply_axes(a.contiguous(), (g.BxT, g.ByT, g.BzT), xform=SsState.xform)


def project_volume(field, SsState):
    """modes = dV·Σ f·B_z B_y B_x = √dV·DCT(f) (spectral_ops.py:94-102)."""
    g = SsState.grid
    return _apply_axes(field.contiguous(), (g.Bx, g.By, g.Bz), scale=g.dV, xform=SsState.xform)


def _latent(Q_buffer, T, T_prev, phys, num):
    model = getattr(phys, "model", None)
    prog = _program(model) if (model is not None and not model.is_constant) else None
    call("sgh_latent_source", dtype_code(Q_buffer.dtype), Q_buffer.numel(), ptr(T.contiguous()),
         ptr(T_prev.contiguous()), float(phys.T_solidus), float(phys.T_liquidus), float(phys.rho),
         float(phys.L_f), float(num.dt), ptr(prog), ptr(Q_buffer), stream_ptr())


def compute_latent_heat_source_grid(Q_buffer, T_full, T_prev_full, phys, num, SsState):
    """Grid-mode latent source −ρL(f_l(T) − f_l(T_prev))/dt (spectral_ops.py:270-289)."""
    _latent(Q_buffer, T_full, T_prev_full, phys, num)


def _box_latent(Q_buffer, a, SsState, phys, num):
    """Q on the box at modes ``a`` against ``T_prev``: one fused entry (three contractions,
    the latent source formed in the last one's epilogue)."""
    fm = SsState.fine_mesh
    a = a.contiguous()
    nz, ny, nx = a.shape
    nzb, nyb, nxb = fm.shape
    model = getattr(phys, "model", None)
    prog = _program(model) if (model is not None and not model.is_constant) else None
    call("sgh_box_latent_source", dtype_code(a.dtype), SsState.xform, nx, ny, nz, ptr(a),
         nxb, nyb, nzb, fm._col(fm.BxT_full, fm.x_start * nx), nx,
         fm._col(fm.ByT_full, fm.y_start * ny), ny, ptr(fm.BzT), ptr(fm.T_prev),
         float(phys.T_solidus), float(phys.T_liquidus), float(phys.rho), float(phys.L_f),
         float(num.dt), ptr(prog), ptr(fm._wbuf), ptr(Q_buffer), stream_ptr())


def compute_latent_heat_source(Q_buffer, phys, num, SsState):
    """Box-mode latent source at ``buffers.a_temp`` against ``T_prev``; zero with no box or
    no history yet (spectral_ops.py:234-267)."""
    fm = SsState.fine_mesh
    if fm is None or fm.T_prev is None:
        Q_buffer.zero_()
        return
    _box_latent(Q_buffer, SsState.buffers.a_temp, SsState, phys, num)


def initialize_latent_heat_if_needed(SsState, a=None):
    """First step only: ``T_prev`` = the box temperature of ``a`` (default
    ``buffers.a_temp``), ``Q_prev`` = 0 (spectral_ops.py:182-195)."""
    fm = SsState.fine_mesh
    if fm is None or fm.T_prev is not None:
        return
    src = SsState.buffers.a_temp if a is None else a
    fm.T_prev = _reconstruct_temperature_box(src, SsState, out=fm._T)
    if fm.Q_prev is None:
        fm._Q.zero_()
        fm.Q_prev = fm._Q


def update_latent_heat_history(SsState, a=None):
    """``T_prev`` = the box temperature of the converged modes ``a`` (default
    ``buffers.a_temp``) (spectral_ops.py:198-207)."""
    fm = SsState.fine_mesh
    if fm is None:
        return
    src = SsState.buffers.a_temp if a is None else a
    if fm.T_prev is None:
        fm.T_prev = fm._T
    _reconstruct_temperature_box(src, SsState, out=fm.T_prev)


def shift_latent_heat_history(SsState, laser_state, num):
    """Move the box history with the box: shift by −v·dt/d_fine in (y, x), ``T_prev`` with
    edge values, ``Q_prev`` with zeros (spectral_ops.py:292-306)."""
    fm = SsState.fine_mesh
    if fm is None or fm.T_prev is None:
        return
    sy = -(laser_state.v[1] * num.dt) / fm.dy_fine
    sx = -(laser_state.v[0] * num.dt) / fm.dx_fine
    nzb, nyb, nxb = fm.shape
    code, st = dtype_code(fm.T_prev.dtype), stream_ptr()
    dst = fm._T_alt if fm.T_prev is fm._T else fm._T
    call("sgh_shift3d_xy", code, nzb, nyb, nxb, ptr(fm.T_prev), float(sy), float(sx), 1,
         ptr(dst), st)
    fm.T_prev = dst
    if fm.Q_prev is not None:
        dst = fm._Q_alt if fm.Q_prev is fm._Q else fm._Q
        call("sgh_shift3d_xy", code, nzb, nyb, nxb, ptr(fm.Q_prev), float(sy), float(sx), 0,
             ptr(dst), st)
        fm.Q_prev = dst


def box_face_peaks(Q_box, SsState):
    """(max |Q|, floor, x faces, y faces) of a box field, one device launch + one read
    (the test of _check_latent_box_truncation, spectral.py:545-586)."""
    fm = SsState.fine_mesh
    nzb, nyb, nxb = fm.shape
    call("sgh_box_face_peaks", dtype_code(Q_box.dtype), nzb, nyb, nxb, ptr(Q_box.contiguous()),
         ptr(fm.peaks), stream_ptr())
    return [float(v) for v in fm.peaks[:4].cpu()]


_PROGRAMS: dict = {}


def _program(model):
    """Packed device property program of a MaterialModel (cached per model)."""
    key = id(model)
    hit = _PROGRAMS.get(key)
    if hit is not None and hit[0] is model:
        return hit[1]
    blob = model.device_program()
    t = torch.frombuffer(bytearray(blob), dtype=torch.uint8).to(_DEVICE)
    _PROGRAMS[key] = (model, t)
    return t


def assemble_property_correction(SsState, a_trial, T_prev_full, dt, model, k_bar, a_bar):
    """C = project_volume(∇·(k'∇T) − a'∂ₜT) at the trial modes (spectral_ops.py:105-170)."""
    g = SsState.grid
    nz, ny, nx = a_trial.shape
    T = reconstruct_volume(a_trial, SsState)
    work = torch.empty((3,) + tuple(a_trial.shape), dtype=a_trial.dtype, device=a_trial.device)
    f = torch.empty_like(T)
    call("sgh_property_correction_source", dtype_code(a_trial.dtype), nx, ny, nz, ptr(T),
         ptr(T_prev_full.contiguous()), ptr(_program(model)), float(k_bar), float(a_bar),
         float(dt), g.dx, g.dy, g.dz, ptr(work), ptr(f), stream_ptr())
    return project_volume(f, SsState)


def shift_flux(field, shift, geom):
    """Translate a surface flux by ``shift = (Δx, Δy)`` metres (spectral_cpu_kernels.py:392-396)."""
    dx, dy = shift
    out = torch.empty_like(field)
    ny, nx = field.shape
    call("sgh_shift2d", dtype_code(field.dtype), ny, nx, ptr(field.contiguous()),
         float(dy) / float(geom.d.y), float(dx) / float(geom.d.x), None, 1.0, ptr(out),
         stream_ptr())
    return out


def laser_flux(SsState, profile, laser_params, x0, y0, peak, out=None):
    """Beam flux on the cell centres (laser.py:42-53,65-90,122-135)."""
    g = SsState.grid
    if out is None:
        out = torch.empty((g.ny, g.nx), dtype=_tdtype(SsState.dtype), device=_DEVICE)
    call("sgh_laser_flux", SsState.code, g.nx, g.ny, g.dx, g.dy, float(x0), float(y0),
         float(laser_params.r_x), float(laser_params.r_y), float(profile.order), float(peak),
         int(bool(profile.cell_integrated)), ptr(out), stream_ptr())
    return out


# ---------------------------------------------------------------------------------
# Fused per-step driver (the plain device form of SpectralSolver.step)
# ---------------------------------------------------------------------------------


class _PlainIter:
    """Descriptor of the native per-iteration entry (``sgh_plain_iterate``) for one solver
    state: every buffer the iteration touches, built once; the box window, the box
    history pointer and the laser-dependent switches are refreshed each step."""

    def __init__(self, solver, S):
        import ctypes
        g, b, fm = S.grid, S.buffers, S.fine_mesh
        mat, num = solver.context.mat, solver.context.num
        td, dev = b.a_temp.dtype, b.a_temp.device
        nx, ny, nz = g.nx, g.ny, g.nz
        shp = (nz, ny, nx)
        d = _lib.SghPlainIter()
        self.d = d
        self.ref = ctypes.byref(d)
        d.dtype, d.xform = S.code, S.xform
        d.nx, d.ny, d.nz, d.nchunks, d.n_pairs = nx, ny, nz, b.nchunks, b.n_pairs
        d.dxdy, d.dV, d.n_elements = g.dx * g.dy, g.dV, float(nx * ny * nz)
        d.Pa, d.T_boil, d.dH = float(mat.Pa), float(mat.T_boil), float(mat.DeltaH_LV)
        d.R_v, d.T_liq, d.T0 = float(mat.R_v), float(mat.T_liquidus), float(mat.T0)
        d.T_S, d.T_L, d.rho = float(mat.T_solidus), float(mat.T_liquidus), float(mat.rho)
        d.L_f, d.dt = float(getattr(mat, "L_f", 0.0)), float(num.dt)
        d.dx, d.dy, d.dz = g.dx, g.dy, g.dz
        d.aK, d.KK = ptr(S.a), ptr(S.KK)
        self.a_buf = (b.a_temp, b.a_alt)
        d.a_buf[0], d.a_buf[1] = ptr(b.a_temp), ptr(b.a_alt)
        d.w_top, d.w_bot = ptr(g.w_top), ptr(g.w_bot)
        d.zc, d.T_s, d.q_las, d.q_tmp = ptr(b.zc), ptr(b.T_s), ptr(b.q_las), ptr(b.q_tmp)
        self.q_evap = (b.q_evap_buffer, torch.zeros_like(b.q_evap_buffer))
        d.q_evap[0], d.q_evap[1] = ptr(self.q_evap[0]), ptr(self.q_evap[1])
        d.S_top, d.S_zero, d.S_bot_raw = ptr(b.S_top), ptr(b.S_zero), ptr(b.S_bot_raw)
        d.sByT, d.sBx, d.sBy, d.sBxT = ptr(g.ByT), ptr(g.Bx), ptr(g.By), ptr(g.BxT)
        self.Q_latent = None
        self.Q_modes = (None, None)
        if fm is not None:
            d.latent_mode = 2
            d.nxb, d.nyb, d.nzb = fm.nx_box, fm.ny_box, fm.nz_box
            d.ntx, d.nty = fm.n_fine_totals[0], fm.n_fine_totals[1]
            d.dV_fine = fm.dV_fine
            d.fBxT, d.fByT, d.fBzT = ptr(fm.BxT_full), ptr(fm.ByT_full), ptr(fm.BzT)
            d.fBx, d.fBy, d.fBz = ptr(fm.Bx_full), ptr(fm.By_full), ptr(fm.Bz)
            d.box_work = ptr(fm._wbuf)
            self.Q_latent = (b.Q_latent_buffer, torch.zeros_like(b.Q_latent_buffer))
            self.Q_modes = (fm.Q_modes, torch.empty_like(fm.Q_modes))
        elif solver._grid_latent:
            d.latent_mode = 1
            self.Q_latent = (b.Q_latent_buffer, torch.zeros_like(b.Q_latent_buffer))
            self.Q_modes = (torch.empty(shp, dtype=td, device=dev),
                            torch.empty(shp, dtype=td, device=dev))
        if self.Q_latent is not None:
            d.Q_latent[0], d.Q_latent[1] = ptr(self.Q_latent[0]), ptr(self.Q_latent[1])
            d.Q_modes[0], d.Q_modes[1] = ptr(self.Q_modes[0]), ptr(self.Q_modes[1])
        d.T_prev_full = ptr(solver._T_prev_full)
        d.vBxT, d.vByT, d.vBzT = ptr(g.BxT), ptr(g.ByT), ptr(g.BzT)
        d.vBx, d.vBy, d.vBz = ptr(g.Bx), ptr(g.By), ptr(g.Bz)
        self.vol = None
        if d.latent_mode == 1 or solver._property_correction:
            self.vol = [torch.empty(shp, dtype=td, device=dev) for _ in range(3)]
            d.vol_w1, d.vol_w2, d.vol_T = (ptr(v) for v in self.vol)
        self.corr = None
        if solver._property_correction:
            self.corr = (torch.empty(shp, dtype=td, device=dev),
                         torch.empty((3,) + shp, dtype=td, device=dev),
                         torch.empty(shp, dtype=td, device=dev),
                         torch.empty(shp, dtype=td, device=dev))
            d.a_raw, d.corr_work, d.corr_f, d.C_modes = (ptr(v) for v in self.corr)
            d.kbar, d.abar = float(solver._kbar), float(solver._abar)
        d.partials, d.norms = ptr(b.partials), ptr(b.norms)
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)
        d.counter = ptr(self.counter)
        # pinned, coherent host words the fused pass writes: per parity the three norms
        # and the sequence number of the iteration that wrote them
        hp = ctypes.c_void_p()
        call("sgh_host_alloc", 8 * 8, ctypes.byref(hp))
        self._host_ptr = hp.value
        self.norms_host = np.ctypeslib.as_array((ctypes.c_double * 8).from_address(hp.value))
        self.norms_host[:] = 0.0
        d.norms_host = hp.value
        self.seq = 0
        self.seq_of = [0, 0]
        # a latent source's chain runs beside the surface chain on a side stream: for a
        # solver that has the GPU to itself and enough modes that the overlap outweighs the
        # two cross-stream waits per iteration (solvers stepped concurrently, as in
        # PlainBatch, turn it off: there the waits stall the hardware queues they share)
        self.forked = False
        if (d.latent_mode != 0 and _PlainIter.FORK and getattr(solver, "_plain_fork", True)
                and nx * ny * nz >= _PlainIter.FORK_MIN_MODES):
            call("sgh_plain_iter_fork_init", self.ref)
            self.forked = True

    # whether new descriptors get a side stream for the surface chain (results are
    # bit-identical either way), and from how many modes (64×32×16 tracks measured slower
    # with it, 128×64×32 and 80×48×120 faster)
    FORK = True
    FORK_MIN_MODES = 100_000

    def __del__(self):
        ptr_ = getattr(self, "_host_ptr", None)
        if ptr_:
            try:
                # a speculative iteration may still be writing these words
                torch.cuda.synchronize()
                if getattr(self, "forked", False):
                    _lib.lib.sgh_plain_iter_fork_release(self.ref)
                    self.forked = False
                _lib.lib.sgh_host_free(ptr_)
            except Exception:  # interpreter shutdown
                pass
            self._host_ptr = None

    # polling of the pinned norm words: a tight spin for the first SPIN_S seconds (an
    # iteration takes tens to hundreds of µs and the next one is already enqueued, so
    # noticing a little late does not idle the GPU), then yields of the host thread, then
    # short sleeps
    SPIN_S, YIELD_S, NAP_S = 2e-4, 2e-3, 5e-5

    def ready(self, k):
        """Whether iteration k's fused pass has written its norms (no wait)."""
        p = k & 1
        return self.norms_host[4 * p + 3] == float(self.seq_of[p])

    def read(self, k):
        """The norms of iteration k (after ``ready(k)``)."""
        h, p = self.norms_host, k & 1
        return float(h[4 * p]), float(h[4 * p + 1]), float(h[4 * p + 2])

    def wait(self, k, timeout_s=60.0):
        """The norms of iteration k once its fused pass has written them."""
        import time as _time
        p = k & 1
        want = float(self.seq_of[p])
        h = self.norms_host
        if h[4 * p + 3] != want:
            t0 = _time.perf_counter()
            while h[4 * p + 3] != want:
                el = _time.perf_counter() - t0
                if el > timeout_s:
                    torch.cuda.synchronize()   # raises a device error if one occurred
                    if h[4 * p + 3] != want:
                        raise _lib.SGHError(f"Picard iteration {k}: no residual after "
                                            f"{timeout_s} s")
                elif el > self.YIELD_S:
                    _time.sleep(self.NAP_S)
                elif el > self.SPIN_S:
                    _time.sleep(0)
        return float(h[4 * p]), float(h[4 * p + 1]), float(h[4 * p + 2])

    def refresh(self, solver, S, prog, kbar, have_bot):
        d, fm = self.d, S.fine_mesh
        d.omega = float(solver.mixing_omega)
        d.h_top = float(solver.context.mat.h_conv_top)
        d.rescale = int(bool(solver._property_correction))
        d.prog = ptr(prog)
        d.kbar = float(kbar)
        d.have_bottom = int(bool(have_bot))
        if fm is not None:
            d.x_start, d.y_start = fm.x_start, fm.y_start
            d.T_prev_box = ptr(fm.T_prev)


def _plain_iter(solver, S, prog, kbar, have_bot):
    pi = getattr(S, "_plain_iter", None)
    if pi is None or pi.d.aK != S.a.data_ptr():
        pi = _PlainIter(solver, S)
        S._plain_iter = pi
    pi.refresh(solver, S, prog, kbar, have_bot)
    return pi


def picard_step(solver, t, dt):
    """One ETD1 step with the relaxed Picard loop, kernel for kernel on the device.

    Follows ``SpectralSolver.step`` (sgh/solvers/spectral.py:219-543): laser flux,
    decay, warm start (shifted evaporation, latent replay, convection), then per
    iteration: surface temperature → evaporation/convection flux → projection →
    [latent and property-correction volume sources] → fused ETD1 + relaxation + FP64
    norms + next-iterate surface contraction, one host read of the residual ratio.
    The host waits for each residual (``picard_step_steps`` is the same step as a
    generator that yields instead, so that many solvers can share the host thread).
    """
    gen = picard_step_steps(solver, t, dt, wait=True)
    while True:
        try:
            next(gen)
        except StopIteration as done:
            return done.value


def picard_step_steps(solver, t, dt, wait=False, timeout_s=60.0):
    """``picard_step`` as a generator: enqueues the step's kernels on the current stream
    and, whenever the next Picard decision needs a residual the device has not written
    yet, yields (``wait=False``) or waits for it (``wait=True``); its return value
    (``StopIteration.value``) is ``(state, metrics)``. A caller that resumes it must make
    the stream it started on current again."""
    import time as _time
    S = solver.state
    g, b = S.grid, S.buffers
    ctx = solver.context
    mat, num, lp = ctx.mat, ctx.num, ctx.laser
    code, st = S.code, stream_ptr()
    nx, ny, nz = g.nx, g.ny, g.nz
    Ssz, N = nx * ny, nx * ny * nz
    dxdy = g.dx * g.dy

    ls = ctx.laser_path.get_state(t, dt)
    vx, vy = ls.v
    peak = (solver._profile.peak_intensity(float(ls.power), float(lp.absorptivity),
                                           float(lp.ref_area)) if ls.is_on else 0.0)
    laser_flux(S, solver._profile, lp, ls.x, ls.y, peak, out=b.q_las)
    call("sgh_reduce", code, Ssz, ptr(b.q_las), 0, ptr(b.scal), st)

    call("sgh_mul_inplace", code, N, ptr(S.a), ptr(S.K), st)
    fm = S.fine_mesh
    if fm is not None:
        # box mode: move the window, seed the history on the first step, shift it
        fm.update(ls)
        initialize_latent_heat_if_needed(S, S.a)
        shift_latent_heat_history(S, ls, num)
    if solver.track_picard_history:
        solver.picard_history = []

    # warm start: S_top = dx·dy·Bᵧ(q_las − shift(q_evap_old))Bₓᵀ
    call("sgh_shift2d", code, ny, nx, ptr(b.q_evap_old), float(vy) * dt / g.dy,
         float(vx) * dt / g.dx, ptr(b.q_las), -1.0, ptr(b.q_tmp), st)
    _sep2d(b.q_tmp, g.By, g.BxT, dxdy, out=b.S_top)
    a_cur = b.a_temp
    Q_warm = None
    if fm is not None and fm.T_prev is not None and fm.Q_prev is not None:
        Q_warm = project_box_to_modes(fm.Q_prev, S, out=fm.Q_modes)
    elif solver._grid_latent and solver._Q_latent_modes_prev is not None:
        Q_warm = solver._Q_latent_modes_prev
    call("sgh_etd1_update", code, nx, ny, nz, ptr(S.a), ptr(S.KK), ptr(g.w_top), ptr(b.S_top),
         None, None, ptr(Q_warm), ptr(a_cur), st)
    h_top, h_bot, T0 = float(mat.h_conv_top), float(mat.h_conv_bottom), float(mat.T0)
    if h_top > 0:
        _recon_face(a_cur, g.w_top, S, out=b.T_s)
        call("sgh_convection_flux", code, Ssz, ptr(b.T_s), h_top, T0, None, 0.0, ptr(b.q_tmp), st)
        _sep2d(b.q_tmp, g.By, g.BxT, dxdy, out=b.S_tmp)
        call("sgh_add_surface_source", code, nx, ny, nz, ptr(a_cur), ptr(S.KK), ptr(g.w_top),
             ptr(b.S_tmp), st)
    if h_bot > 0:
        _recon_face(a_cur, g.w_bot, S, out=b.T_s)
        call("sgh_convection_flux", code, Ssz, ptr(b.T_s), h_bot, T0, None, 0.0, ptr(b.q_tmp), st)
        _sep2d(b.q_tmp, g.By, g.BxT, dxdy, out=b.S_bot)
        call("sgh_add_surface_source", code, nx, ny, nz, ptr(a_cur), ptr(S.KK), ptr(g.w_bot),
             ptr(b.S_bot), st)

    rescale = solver._property_correction
    prog = _program(mat.model) if rescale else None
    kbar = float(solver._kbar) if rescale else 0.0
    have_bot = h_bot > 0
    if have_bot:
        _recon_face(a_cur, g.w_bot, S, out=b.T_s)
        call("sgh_convection_flux", code, Ssz, ptr(b.T_s), h_bot, T0, ptr(prog), kbar,
             ptr(b.q_tmp), st)
        _sep2d(b.q_tmp, g.By, g.BxT, dxdy, out=b.S_bot_raw)

    call("sgh_zcontract", code, nx, ny, nz, ptr(a_cur), ptr(g.w_top), b.nchunks, ptr(b.zc), st)
    tol = float(solver.convergence_tol)
    iter_k, rel, conv = -1, float("nan"), False
    pi = _plain_iter(solver, S, prog, kbar, have_bot)
    maxit = int(solver.max_picard_iter)

    def enqueue(k):
        pi.seq += 1
        pi.seq_of[k & 1] = pi.seq
        pi.d.seq = float(pi.seq)
        call("sgh_plain_iterate", pi.ref, k & 1, st)

    # The next iteration is enqueued before this one's residual is read, so the device
    # never waits on the host; when iteration k converges, iteration k + 1 (which wrote
    # only the other half of every alternating buffer) is discarded.
    if maxit > 0:
        enqueue(0)
    for iter_k in range(maxit):
        if iter_k + 1 < maxit:
            enqueue(iter_k + 1)
        if wait:
            rms_diff, _, rel = pi.wait(iter_k, timeout_s)
        else:
            t_wait = None
            while not pi.ready(iter_k):
                if t_wait is None:
                    t_wait = _time.perf_counter()
                elif _time.perf_counter() - t_wait > timeout_s:
                    torch.cuda.synchronize()   # raises a device error if one occurred
                    if not pi.ready(iter_k):
                        raise _lib.SGHError(f"Picard iteration {iter_k}: no residual after "
                                            f"{timeout_s} s")
                yield
            rms_diff, _, rel = pi.read(iter_k)
        if solver.track_picard_history:
            rho_k = None
            if iter_k > 0 and solver.picard_history:
                prev = solver.picard_history[-1]["rms_diff"]
                if prev > 0:
                    rho_k = rms_diff / prev
            solver.picard_history.append({"iter": iter_k, "rms_diff": rms_diff, "rho": rho_k,
                                          "true_rel_err": rel})
        if rel < tol:
            conv = True
            break
    else:
        if maxit > 0:
            solver._note_cap(t, rel)
    last = iter_k & 1 if iter_k >= 0 else None
    if last is not None:
        a_cur = pi.a_buf[last ^ 1]
        b.q_evap_buffer = pi.q_evap[last]
        if pi.Q_latent is not None:
            b.Q_latent_buffer = pi.Q_latent[last]

    T_temp = _recon_face(a_cur, g.w_top, S, out=b.T_s)
    call("sgh_reduce", code, Ssz, ptr(T_temp), 1, ptr(b.scal[1:]), st)
    if solver._grid_latent and last is not None:
        solver._Q_latent_modes_prev = pi.Q_modes[last]
    b.q_evap_old.copy_(b.q_evap_buffer)
    S.a.copy_(a_cur)
    if solver._property_correction or solver._grid_latent:
        solver._T_prev_full.copy_(reconstruct_volume(S.a, S))
    if fm is not None:
        update_latent_heat_history(S, S.a)
        if fm.Q_prev is None:
            fm._Q.zero_()
            fm.Q_prev = fm._Q
        if last is not None and fm.T_prev is not None:
            fm.Q_prev.copy_(b.Q_latent_buffer)
            _check_box_truncation(solver, S, b.Q_latent_buffer)
    metrics = {
        "T_surface_max": b.scal[1].clone(),
        "P_laser": b.scal[0] * dxdy,
        "n_evap_iter": iter_k + 1,
        "picard_converged": conv,
        "picard_rel_err": rel,
    }
    return S, metrics


def _check_box_truncation(solver, S, Q_box):
    """Warn once per run when the latent source reaches the box floor or a sliding window
    edge (> 1e-2 of its peak): the pool has outgrown the box (spectral.py:545-586)."""
    if getattr(solver, "_latent_truncation_warned", False):
        return
    fm = S.fine_mesh
    peak, floor_, xf, yf = box_face_peaks(Q_box, S)
    if peak <= 0.0:
        return
    faces = {"floor (depth)": floor_}
    if fm.nx_box < fm.n_fine_totals[0]:
        faces["x window edge"] = xf
    if fm.ny_box < fm.n_fine_totals[1]:
        faces["y window edge"] = yf
    hit = [name for name, v in faces.items() if v > 1e-2 * peak]
    if not hit:
        return
    import logging
    logging.getLogger("spectral_galerkin_heat_amd.solvers.spectral").warning(
        "Latent-heat source reaches the fine-mesh box boundary (%s): the melt pool extends "
        "beyond the refined sub-box, so its latent heat is truncated. Enlarge "
        "fine_mesh.box_size (currently %.3g x %.3g x %.3g mm) to enclose the pool.",
        ", ".join(hit), fm.nx_box * fm.dx_fine * 1e3, fm.ny_box * fm.dy_fine * 1e3,
        fm.nz_box * fm.dz_fine * 1e3)
    solver._latent_truncation_warned = True

"""Spectral ETD1 + relaxed-Picard solver on an injected ``MathBackend``.

Mirrors ``SpectralSolver`` (sgh/solvers/spectral.py:38-628): the same constructor
arguments and Picard-parameter precedence (explicit > config > defaults ω = 0.1,
tol = 1e-4, cap = 30; ``_resolve_picard_params`` :107-123), the same initial condition
(``a[0,0,0] = T0·√(LxLyLz)``, :176-180), grid-mode / box-mode latent and
property-correction switches (:185-195), the once-per-run box truncation warning
(:545-586), ``set_state`` (:588-619) and the metrics dictionary of ``step``
(:536-543). The step itself runs through ``backend.kernels.picard_step`` -- on the
MI355X backend, a sequence of fused HIP kernels with one residual read per iteration.
"""

from __future__ import annotations

import logging
import math
from typing import Any

import torch

from ..core.laser import build_laser_profile
from .base import HeatSolver

logger = logging.getLogger(__name__)

__all__ = ["SpectralSolver"]


class SpectralSolver(HeatSolver):
    def __init__(self, backend, context=None, *, mixing_omega=None, convergence_tol=None,
                 max_picard_iter=None):
        self.backend = backend
        self.context = context
        self.state = None
        self._omega_req = mixing_omega
        self._tol_req = convergence_tol
        self._maxit_req = max_picard_iter
        import numpy as np
        self._resolve_picard_params(np.float32)
        self.track_picard_history = False
        self.picard_history = []
        self._property_correction = False
        self._grid_latent = False
        self._T_prev_full = None
        self._kbar = None
        self._abar = None
        self._picard_cap_warned = False
        self._latent_truncation_warned = False
        self._Q_latent_modes_prev = None
        self._profile = build_laser_profile("gaussian")

    def _resolve_picard_params(self, dtype) -> None:
        num = self.context.num if self.context is not None else None

        def pick(explicit, attr, default):
            if explicit is not None:
                return explicit
            if num is not None and getattr(num, attr, None) is not None:
                return getattr(num, attr)
            return default

        self.mixing_omega = dtype(pick(self._omega_req, "picard_omega", 0.1))
        self.convergence_tol = dtype(pick(self._tol_req, "picard_tol", 1e-4))
        self.max_picard_iter = int(pick(self._maxit_req, "max_picard_iter", 30))

    def initialize(self, context=None) -> Any:
        if context is not None:
            self.context = context
        if self.context is None:
            raise RuntimeError("SimulationContext must be provided either at construction "
                               "or in initialize().")
        kernels = self.backend.kernels
        geom, num, mat = self.context.geom, self.context.num, self.context.mat
        self.state = kernels.SpectralSolverState(mat, geom, num, self.context.fine)
        dtype = self.state.dtype
        laser = self.context.laser
        self._profile = build_laser_profile(laser.profile, laser.super_gaussian_order,
                                            cell_integrated=laser.cell_integrated)
        self._profile.validate(geom.d.x, geom.d.y)
        self._resolve_picard_params(dtype)

        xp = self.backend.xp
        self.state.a = xp.zeros((num.nz, num.ny, num.nx), dtype=dtype)
        T0 = mat.T0
        self.state.a[0, 0, 0] = float(dtype(T0 * math.sqrt(geom.size.x * geom.size.y * geom.size.z)))

        self._grid_latent = (self.state.fine_mesh is None and float(getattr(mat, "L_f", 0.0)) > 0.0)
        model = getattr(mat, "model", None)
        self._property_correction = model is not None and not model.is_constant
        if self._property_correction:
            self._kbar = dtype(float(mat.k))
            self._abar = dtype(float(mat.rho) * float(mat.Cp))
        if self._property_correction or self._grid_latent:
            self._T_prev_full = xp.full((num.nz, num.ny, num.nx), dtype(T0), dtype=dtype)
        self._Q_latent_modes_prev = None
        self._latent_truncation_warned = False
        if self._property_correction and self._grid_latent:
            logger.warning(
                "Temperature-dependent run in grid mode (no fine_mesh): the sharp latent-heat "
                "source is projected on the coarse grid and can ring.")
        return self.state

    def _note_cap(self, t, rel) -> None:
        if not self._picard_cap_warned:
            self._picard_cap_warned = True
            logger.warning(
                "Picard iteration hit max_picard_iter=%d at t=%.6g s without reaching "
                "picard_tol=%g (last relative residual %.3g). The returned field is not "
                "converged. Warned once per run.", self.max_picard_iter, t,
                float(self.convergence_tol), rel)

    def step(self, t: float, dt: float):
        if self.state is None:
            raise RuntimeError("Solver must be initialized before calling step().")
        return self.backend.kernels.picard_step(self, t, dt)

    def set_state(self, temperature_field) -> None:
        if self.state is None or self.context is None:
            raise RuntimeError("Solver must be initialized before calling set_state().")
        xp = self.backend.xp
        kernels = self.backend.kernels
        geom = self.context.geom
        dtype = self.state.dtype
        T = xp.asarray(temperature_field, dtype=dtype).contiguous()
        scale = float(dtype(math.sqrt(float(geom.d.x * geom.d.y * geom.d.z))))
        self.state.a = (kernels.DCT_II(T) * scale).contiguous()

    # -- full nonlinear checkpoint --------------------------------------------------------
    CHECKPOINT_VERSION = 1

    def checkpoint(self, step: int = -1) -> dict:
        """Everything a later step reads besides the configuration, as host arrays.

        ``set_state`` (spectral.py:588-619) restores only the modes; a nonlinear run also
        carries the previous evaporation flux, the previous full-volume temperature
        (property correction / grid-mode latent heat), the replayed latent modes (grid
        mode) and the box history (box mode). Restoring all of them makes a resumed run
        repeat the uninterrupted one bit for bit.
        """
        if self.state is None:
            raise RuntimeError("Solver must be initialized before calling checkpoint().")
        S, b = self.state, self.state.buffers
        host = lambda t: t.detach().cpu().numpy().copy()  # noqa: E731
        ck = {"version": self.CHECKPOINT_VERSION, "step": int(step),
              "shape": list(S.a.shape), "dtype": str(S.a.dtype),
              "a": host(S.a), "q_evap_old": host(b.q_evap_old),
              "picard_cap_warned": bool(self._picard_cap_warned),
              "latent_truncation_warned": bool(self._latent_truncation_warned)}
        if self._T_prev_full is not None:
            ck["T_prev_full"] = host(self._T_prev_full)
        if self._Q_latent_modes_prev is not None:
            ck["Q_latent_modes_prev"] = host(self._Q_latent_modes_prev)
        fm = S.fine_mesh
        if fm is not None:
            ck["box_shape"] = list(fm.shape)
            if fm.T_prev is not None:
                ck["box_T_prev"] = host(fm.T_prev)
            if fm.Q_prev is not None:
                ck["box_Q_prev"] = host(fm.Q_prev)
        return ck

    def restore(self, ck: dict) -> int:
        """Load a ``checkpoint()`` into an initialized solver of the same configuration;
        returns the step it was taken after (-1 if none was recorded)."""
        if self.state is None:
            raise RuntimeError("Solver must be initialized before calling restore().")
        S, b = self.state, self.state.buffers
        if int(ck.get("version", -1)) != self.CHECKPOINT_VERSION:
            raise ValueError(f"checkpoint version {ck.get('version')} is not "
                             f"{self.CHECKPOINT_VERSION}")
        if list(ck["shape"]) != list(S.a.shape) or str(ck["dtype"]) != str(S.a.dtype):
            raise ValueError(f"checkpoint of shape {list(ck['shape'])} / {ck['dtype']} does not "
                             f"match this solver ({list(S.a.shape)} / {S.a.dtype})")

        def put(dst, arr):
            dst.copy_(torch.as_tensor(arr, dtype=dst.dtype))

        put(S.a, ck["a"])
        put(b.q_evap_old, ck["q_evap_old"])
        if "T_prev_full" in ck:
            if self._T_prev_full is None:
                raise ValueError("checkpoint carries T_prev_full; this solver has no "
                                 "property correction or grid-mode latent heat")
            put(self._T_prev_full, ck["T_prev_full"])
        self._Q_latent_modes_prev = None
        if "Q_latent_modes_prev" in ck:
            self._Q_latent_modes_prev = torch.as_tensor(ck["Q_latent_modes_prev"]).to(
                device=S.a.device, dtype=S.a.dtype).contiguous()
        fm = S.fine_mesh
        if ("box_shape" in ck) != (fm is not None):
            raise ValueError("checkpoint and solver disagree on box mode")
        if fm is not None:
            if list(ck["box_shape"]) != list(fm.shape):
                raise ValueError(f"checkpoint box {list(ck['box_shape'])} != {list(fm.shape)}")
            fm.T_prev = fm.Q_prev = None
            if "box_T_prev" in ck:
                put(fm._T, ck["box_T_prev"])
                fm.T_prev = fm._T
            if "box_Q_prev" in ck:
                put(fm._Q, ck["box_Q_prev"])
                fm.Q_prev = fm._Q
        self._picard_cap_warned = bool(ck.get("picard_cap_warned", False))
        self._latent_truncation_warned = bool(ck.get("latent_truncation_warned", False))
        if torch.cuda.is_available():
            torch.cuda.synchronize()
        return int(ck.get("step", -1))

    def save_checkpoint(self, path, step: int = -1) -> None:
        """``checkpoint()`` as an uncompressed ``.npz`` (arrays and scalars only, no pickle)."""
        import numpy as np
        ck = self.checkpoint(step)
        flat = {k: (np.asarray(v) if not isinstance(v, str) else np.array(v)) for k, v in ck.items()}
        np.savez(path, **flat)

    def load_checkpoint(self, path) -> int:
        import numpy as np
        with np.load(path, allow_pickle=False) as z:
            ck = {k: z[k] for k in z.files}
        for k in ("version", "step"):
            ck[k] = int(ck[k])
        for k in ("shape", "box_shape"):
            if k in ck:
                ck[k] = [int(v) for v in ck[k]]
        ck["dtype"] = str(ck["dtype"])
        for k in ("picard_cap_warned", "latent_truncation_warned"):
            if k in ck:
                ck[k] = bool(ck[k])
        return self.restore(ck)

    def finalize(self) -> None:
        """Release device buffers held by the state."""
        self.state = None
        if torch.cuda.is_available():
            torch.cuda.synchronize()

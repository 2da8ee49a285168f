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

    def finalize(self) -> None:
        """Release device buffers held by the state."""
        self.state = None
        if torch.cuda.is_available():
            torch.cuda.synchronize()

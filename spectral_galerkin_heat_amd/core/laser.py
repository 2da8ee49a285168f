"""Laser beam profiles and trajectories.

Mirrors ``src/spectral_galerkin_heat/core/laser.py``: the super-Gaussian area factor
(:34-40), ``LaserProfile`` with the energy-conserving peak ``I₀ = A·P/(f·r_x·r_y)``
(:93-135), ``build_laser_profile`` (:148-165), ``LaserState`` and the abstract
``LaserPath`` (:168-217). The flux field itself is evaluated on the device by the
surface-forcing kernel; this module only resolves the profile parameters.

Also holds the concrete paths the parity and benchmark cases use: a fixed beam and a
constant-velocity track (the reference's test fixtures,
``tests/unit_tests/conftest.py:11-18`` and ``tests/integration/conftest.py:34-54``) and a
multi-track raster with laser-off dwells (not expressible in G-code, which has no dwell).
"""

from __future__ import annotations

import math
from abc import ABC, abstractmethod
from dataclasses import dataclass

__all__ = [
    "ConstantVelocityLaser", "FixedLaser", "LaserPath", "LaserProfile", "LaserState",
    "RasterLaserPath", "SingleTrackLaser", "build_laser_profile",
    "super_gaussian_area_factor",
]


def super_gaussian_area_factor(order: float) -> float:
    """``f(n) = π·2^(1−2/n)·Γ(2/n)/n`` so that ``∬ exp(−2ρⁿ) dA = f·r_x·r_y``."""
    n = float(order)
    return math.pi * 2.0 ** (1.0 - 2.0 / n) * math.gamma(2.0 / n) / n


@dataclass(frozen=True)
class LaserProfile:
    name: str
    order: float
    cell_integrated: bool = False

    @property
    def area_factor(self) -> float:
        return super_gaussian_area_factor(self.order)

    def peak_intensity(self, power: float, absorptivity: float, ref_area: float) -> float:
        return absorptivity * power / (self.area_factor * ref_area)

    def validate(self, hx=None, hy=None) -> None:
        if self.cell_integrated:
            if self.order != 2.0:
                raise ValueError(
                    "cell_integrated flux is only defined for the Gaussian "
                    f"(order 2), got order {self.order}.")
            if hx is None or hy is None:
                raise ValueError("cell_integrated flux requires cell sizes hx, hy.")


_PROFILE_ORDERS = {"gaussian": 2.0, "flat_top": 12.0, "super_gaussian": None}


def build_laser_profile(name: str, order: float = 2.0,
                        cell_integrated: bool = False) -> LaserProfile:
    key = str(name).lower()
    try:
        canonical = _PROFILE_ORDERS[key]
    except KeyError:
        choices = ", ".join(sorted(_PROFILE_ORDERS))
        raise ValueError(f"Unknown laser profile: {name!r}. Choose one of: {choices}.") from None
    resolved = float(order) if canonical is None else canonical
    return LaserProfile(name=key, order=resolved, cell_integrated=bool(cell_integrated))


@dataclass
class LaserState:
    """Laser snapshot: position (m), power (W), on flag, velocity (m/s)."""

    x: float
    y: float
    power: float
    is_on: bool
    v: tuple = (0.0, 0.0)


class LaserPath(ABC):
    """Position and power as a function of time; ``get_state`` is called once per step."""

    @abstractmethod
    def get_state(self, time: float, dt: float) -> LaserState:
        ...


class FixedLaser(LaserPath):
    """Stationary beam at ``(x, y)`` with constant power."""

    def __init__(self, x, y, power, is_on=True):
        self._x, self._y, self._power, self._on = x, y, power, is_on

    def get_state(self, time, dt):
        return LaserState(x=self._x, y=self._y, power=self._power, is_on=self._on)


class ConstantVelocityLaser(LaserPath):
    """Straight line from ``(x0, y0)`` at velocity ``(vx, vy)``; always on unless told."""

    def __init__(self, x0, y0, vx, vy, power, is_on=True):
        self._x0, self._y0, self._vx, self._vy = x0, y0, vx, vy
        self._power, self._on = power, is_on

    def get_state(self, time, dt):
        return LaserState(x=self._x0 + self._vx * time, y=self._y0 + self._vy * time,
                          power=self._power, is_on=self._on, v=(self._vx, self._vy))


class SingleTrackLaser(LaserPath):
    """One x-track from ``x_start`` to ``x_end`` at speed ``v``; off (at rest) afterwards."""

    def __init__(self, x_start, x_end, y, speed, power):
        self.x_start, self.x_end, self.y = float(x_start), float(x_end), float(y)
        self.speed, self.power = float(speed), float(power)
        self.t_end = abs(self.x_end - self.x_start) / self.speed
        self.dir = 1.0 if self.x_end >= self.x_start else -1.0

    def get_state(self, time, dt):
        if time <= self.t_end:
            return LaserState(x=self.x_start + self.dir * self.speed * time, y=self.y,
                              power=self.power, is_on=True, v=(self.dir * self.speed, 0.0))
        return LaserState(x=self.x_end, y=self.y, power=0.0, is_on=False, v=(0.0, 0.0))


class RasterLaserPath(LaserPath):
    """Parallel x-tracks with alternating direction and a laser-off dwell between them.

    Track ``k`` runs along y = ``y_center + (k − (n−1)/2)·hatch`` from ``x_start`` to
    ``x_end`` (reversed on odd tracks) at ``speed``; the beam is off and at rest for
    ``dwell`` seconds after each track, and off after the last one.
    """

    def __init__(self, n_tracks, x_start, x_end, y_center, hatch, speed, power, dwell):
        self.n = int(n_tracks)
        self.x_start, self.x_end = float(x_start), float(x_end)
        self.y_center, self.hatch = float(y_center), float(hatch)
        self.speed, self.power, self.dwell = float(speed), float(power), float(dwell)
        self.t_track = abs(self.x_end - self.x_start) / self.speed
        self.period = self.t_track + self.dwell

    def _track_y(self, k):
        return self.y_center + (k - 0.5 * (self.n - 1)) * self.hatch

    def get_state(self, time, dt):
        k = int(math.floor(time / self.period)) if time >= 0 else 0
        if k >= self.n:
            k = self.n - 1
            fwd = (k % 2 == 0)
            x = self.x_end if fwd else self.x_start
            return LaserState(x=x, y=self._track_y(k), power=0.0, is_on=False, v=(0.0, 0.0))
        tau = time - k * self.period
        fwd = (k % 2 == 0)
        x0, x1 = (self.x_start, self.x_end) if fwd else (self.x_end, self.x_start)
        sgn = 1.0 if x1 >= x0 else -1.0
        if tau <= self.t_track:
            return LaserState(x=x0 + sgn * self.speed * tau, y=self._track_y(k),
                              power=self.power, is_on=True, v=(sgn * self.speed, 0.0))
        return LaserState(x=x1, y=self._track_y(k), power=0.0, is_on=False, v=(0.0, 0.0))

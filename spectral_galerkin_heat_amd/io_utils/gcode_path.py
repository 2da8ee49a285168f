"""Laser path read from a G-code file, held as arrays of timed segments.

The semantics are those of the reference's ``GCodeLaserPath``
(``src/spectral_galerkin_heat/io_utils/gcode_path.py:15-119``), pinned by its unit tests
(``tests/unit_tests/io_utils/test_gcode_path.py:24-58``) and its example paths:

* recognised lines (after stripping; empty lines and lines starting with ``;`` are
  skipped): ``G21`` (millimetres) and ``G20`` (inches) set the scale to metres; a line
  starting with ``G0`` or ``G1`` is a move whose ``X``/``Y``/``F`` words (every
  whitespace-separated word starting with that letter) set the target and the feed, each
  scaled and rounded to float32; ``M3`` switches the beam on with the power of its ``S``
  word, ``M5`` switches it off and zeroes the power;
* a move made while the beam is on becomes a segment lasting ``distance / (feed / 60)``
  (1 m/s when no feed was given); moves with the beam off take no time, so the clock runs
  only along lit segments;
* at time t inside a segment the beam is on at the float32 linear interpolation, with the
  finite-difference velocity over ``dt`` (zero at the segment start); after the last
  segment it is off at the path's end, before the first off at ``initial_position``
  (scaled by the file's unit). Segment times and points are float32; the query time
  enters the comparisons and the interpolation as given, so a Python float is compared
  and interpolated in float32 and a float64 scalar in float64.

The file is parsed once into NumPy arrays (start and end points, start and end times,
power); a query finds its segment by binary search over the end times, which rise with
the segment index.
"""

from __future__ import annotations

import math

import numpy as np

from ..core.laser import LaserPath, LaserState

__all__ = ["GCodeLaserPath", "read_gcode_segments"]

_SCALE_TO_M = {"G21": 1.0 / 1000.0, "G20": 25.4 / 1000.0}
_F32_ZERO = np.float32(0.0)


def _last_value(words, letter, scale=1.0):
    """Value of the last word that starts with ``letter``, times ``scale`` (None if no word
    does). Every such word must be a number."""
    values = [float(w[1:]) * scale for w in words if w.startswith(letter)]
    return values[-1] if values else None


def read_gcode_segments(path):
    """Parse a G-code file into ``(table, unit_scale)``.

    ``table`` is a dict of float32 arrays: ``p0`` and ``p1`` (segments × 2, start and end
    point in metres), ``t0`` and ``t1`` (start and end time in seconds) and ``power`` (W).
    """
    scale = 1.0
    xy = [0.0, 0.0]                 # current point (float32 values held as Python floats)
    prev = None                     # point of the previous move
    feed = np.float32(0.0)
    power, lit = 0.0, False
    clock = 0.0                     # seconds along the lit segments (float64)
    rows = []
    with open(path) as fh:
        for raw in fh:
            line = raw.strip()
            if not line or line.startswith(";"):
                continue
            code = line[:3]
            if code in _SCALE_TO_M:
                scale = _SCALE_TO_M[code]
                continue
            words = line.split()
            if line.startswith(("G0", "G1")):
                for axis, letter in enumerate("XY"):
                    v = _last_value(words, letter, scale)
                    if v is not None:
                        xy[axis] = float(np.float32(v))
                v = _last_value(words, "F", scale)
                if v is not None:
                    feed = np.float32(v)
                if prev is not None and lit:
                    ddx, ddy = xy[0] - prev[0], xy[1] - prev[1]
                    speed = float(feed) / 60.0 if feed > 0 else 1.0
                    t_end = clock + math.sqrt(ddx * ddx + ddy * ddy) / speed
                    rows.append((prev[0], prev[1], xy[0], xy[1], clock, t_end, power))
                    clock = t_end
                prev = (xy[0], xy[1])
            elif line.startswith("M3"):
                v = _last_value(words, "S")
                if v is not None:
                    power = v
                lit = True
            elif line.startswith("M5"):
                lit, power = False, 0.0
    arr = np.array(rows, dtype=np.float64).reshape(-1, 7).astype(np.float32)
    table = {"p0": arr[:, 0:2].copy(), "p1": arr[:, 2:4].copy(), "t0": arr[:, 4].copy(),
             "t1": arr[:, 5].copy(), "power": arr[:, 6].copy()}
    return table, scale


class GCodeLaserPath(LaserPath):
    """``LaserPath`` of a G-code file (lit moves become timed segments)."""

    def __init__(self, gcode_file, initial_position=(0.0, 0.0)):
        self._seg, self.unit_scale = read_gcode_segments(gcode_file)
        s32 = np.float32(self.unit_scale)
        self.initial_position = (np.float32(initial_position[0]) * s32,
                                 np.float32(initial_position[1]) * s32)

    @property
    def n_segments(self) -> int:
        return int(self._seg["t0"].shape[0])

    @property
    def segments(self):
        """Segments as ``((x0, y0), (x1, y1), t0, t1, power, is_on)`` tuples of float32."""
        s = self._seg
        return [((s["p0"][i, 0], s["p0"][i, 1]), (s["p1"][i, 0], s["p1"][i, 1]), s["t0"][i],
                 s["t1"][i], s["power"][i], True) for i in range(self.n_segments)]

    @staticmethod
    def _off(x, y) -> LaserState:
        return LaserState(x=np.float32(x), y=np.float32(y), power=_F32_ZERO, is_on=False,
                          v=(_F32_ZERO, _F32_ZERO))

    def get_state(self, time: float, dt: float) -> LaserState:
        # ``time`` takes part in every comparison and product as given, so NumPy's scalar
        # promotion sets the arithmetic: float32 for a Python float, float64 for np.float64
        s = self._seg
        n = self.n_segments
        if n:
            kind = np.result_type(s["t1"], time)
            k = int(np.searchsorted(s["t1"].astype(kind), np.asarray(time, dtype=kind),
                                    side="left"))  # first segment ending at or after time
        else:
            k = 0
        if k < n and s["t0"][k] <= time:
            t0, t1 = s["t0"][k], s["t1"][k]
            p0, d = s["p0"][k], s["p1"][k] - s["p0"][k]
            frac = (time - t0) / (t1 - t0) if t1 > t0 else 0.0
            here = (p0 + frac * d).astype(np.float32)
            vel = (_F32_ZERO, _F32_ZERO)
            if time > t0 and dt > 0.0:
                back = time - dt
                f_back = (back - t0) / (t1 - t0) if (back >= t0 and t1 - t0 > 0) else 0.0
                there = p0 + f_back * d
                vel = tuple(np.float32((float(here[i]) - float(there[i])) / float(dt))
                            for i in range(2))
            return LaserState(x=here[0], y=here[1], power=s["power"][k], is_on=True, v=vel)
        if n and time > s["t1"][-1]:
            return self._off(s["p1"][-1, 0], s["p1"][-1, 1])
        return self._off(*self.initial_position)

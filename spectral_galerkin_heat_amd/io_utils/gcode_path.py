"""G-code laser path: G0/G1 moves, M3 S<power> / M5, G20/G21 units.

Restates ``src/spectral_galerkin_heat/io_utils/gcode_path.py:15-119``: positions are
quantised to float32 after unit scaling, a segment is emitted only for a move made
while the laser is on, off-moves take no time, and the state interpolates linearly in
time along the active segment with a finite-difference velocity over ``dt``. Before
the first segment the beam is off at ``initial_position``; after the last it is off
at the final point.
"""

from __future__ import annotations

import numpy as np

from ..core.laser import LaserPath, LaserState

__all__ = ["GCodeLaserPath"]


class GCodeLaserPath(LaserPath):
    def __init__(self, gcode_file, initial_position=(0.0, 0.0)):
        self.segments, self.unit_scale = self._parse_gcode(gcode_file)
        self.initial_position = (
            np.float32(initial_position[0]) * np.float32(self.unit_scale),
            np.float32(initial_position[1]) * np.float32(self.unit_scale),
        )
        self.current_segment = 0

    @staticmethod
    def _parse_gcode(filepath):
        segments = []
        pos = [0.0, 0.0]
        power, on, feed = 0.0, False, 0.0
        last_pos, last_t, scale = None, 0.0, 1.0
        with open(filepath) as f:
            for raw in f:
                line = raw.strip()
                if not line or line.startswith(";"):
                    continue
                if line.startswith("G21"):
                    scale = 1.0 / 1000.0
                    continue
                if line.startswith("G20"):
                    scale = 25.4 / 1000.0
                    continue
                if line.startswith(("G0", "G1")):
                    x = y = None
                    for tok in line.split():
                        if tok.startswith("X"):
                            x = np.float32(float(tok[1:]) * scale)
                        elif tok.startswith("Y"):
                            y = np.float32(float(tok[1:]) * scale)
                        elif tok.startswith("F"):
                            feed = np.float32(float(tok[1:]) * scale)
                    if x is not None:
                        pos[0] = float(x)
                    if y is not None:
                        pos[1] = float(y)
                    if last_pos is not None and on:
                        ddx, ddy = pos[0] - last_pos[0], pos[1] - last_pos[1]
                        dist = float(np.sqrt(ddx * ddx + ddy * ddy))
                        speed = float(feed) / 60.0 if feed > 0 else 1.0
                        seg_dt = dist / speed if speed > 0 else 0.0
                        t1 = float(last_t + seg_dt)
                        segments.append((
                            (np.float32(last_pos[0]), np.float32(last_pos[1])),
                            (np.float32(pos[0]), np.float32(pos[1])),
                            np.float32(last_t), np.float32(t1), np.float32(power), bool(on)))
                        last_t = t1
                    last_pos = list(pos)
                elif line.startswith("M3"):
                    for tok in line.split():
                        if tok.startswith("S"):
                            power = float(tok[1:])
                    on = True
                elif line.startswith("M5"):
                    on = False
                    power = 0.0
        return segments, scale

    def get_state(self, time: float, dt: float) -> LaserState:
        for start, end, t0, t1, power, on in self.segments:
            if t0 <= time <= t1:
                frac = (time - t0) / (t1 - t0) if t1 > t0 else 0.0
                x = np.float32(start[0] + frac * (end[0] - start[0]))
                y = np.float32(start[1] + frac * (end[1] - start[1]))
                if time > t0 and dt > 0.0:
                    pf = (((time - dt) - t0) / (t1 - t0)
                          if (time - dt) >= t0 and (t1 - t0) > 0 else 0.0)
                    px = start[0] + pf * (end[0] - start[0])
                    py = start[1] + pf * (end[1] - start[1])
                    vx = np.float32((float(x) - float(px)) / float(dt))
                    vy = np.float32((float(y) - float(py)) / float(dt))
                else:
                    vx = vy = np.float32(0.0)
                return LaserState(x=np.float32(x), y=np.float32(y), power=np.float32(power),
                                  is_on=on, v=(vx, vy))
        if self.segments and time > self.segments[-1][3]:
            end = self.segments[-1][1]
            return LaserState(x=np.float32(end[0]), y=np.float32(end[1]),
                              power=np.float32(0.0), is_on=False,
                              v=(np.float32(0.0), np.float32(0.0)))
        return LaserState(x=np.float32(self.initial_position[0]),
                          y=np.float32(self.initial_position[1]), power=np.float32(0.0),
                          is_on=False, v=(np.float32(0.0), np.float32(0.0)))

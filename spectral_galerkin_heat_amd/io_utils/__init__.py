"""Host-side inputs to the step: G-code laser paths."""

from .gcode_path import GCodeLaserPath

__all__ = ["GCodeLaserPath"]

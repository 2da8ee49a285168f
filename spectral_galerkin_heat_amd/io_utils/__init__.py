"""Host-side inputs to the step (G-code laser paths) and output-side error norms."""

from .gcode_path import GCodeLaserPath
from .l2_error import compute_L2_structured, trapezoidal_weights

__all__ = ["GCodeLaserPath", "compute_L2_structured", "trapezoidal_weights"]

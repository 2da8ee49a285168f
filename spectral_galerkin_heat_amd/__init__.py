"""MI355X-native backend for the spectral Galerkin ETD1 + Picard heat solver.

Public names mirror the reference package (``src/spectral_galerkin_heat/__init__.py``):
the configuration layer, the backend registry with an ``"mi355x"`` backend, and the
``SpectralSolver`` / ``build_solver`` entry points. Solvers and backends load lazily so
that importing the configuration layer never touches the GPU library.
"""

__version__ = "0.1.0"

from .core import (GeomParams, LaserParams, LaserPath, LaserState, MaterialParams, NumParams,
                   SimulationContext)

_LAZY = {
    "HeatSolver": (".solvers.base", "HeatSolver"),
    "SpectralSolver": (".solvers.spectral", "SpectralSolver"),
    "build_solver": (".solvers", "build_solver"),
    "MathBackend": (".backends", "MathBackend"),
    "MI355XBackend": (".backends", "MI355XBackend"),
    "get_backend": (".backends", "get_backend"),
    "register_backend": (".backends", "register_backend"),
    "BatchEngine": (".engine.batch", "BatchEngine"),
}


def __getattr__(name):
    if name in _LAZY:
        import importlib
        mod, attr = _LAZY[name]
        obj = getattr(importlib.import_module(mod, package=__name__), attr)
        globals()[name] = obj
        return obj
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")


__all__ = [
    "GeomParams", "LaserParams", "LaserPath", "LaserState", "MaterialParams", "NumParams",
    "SimulationContext", *_LAZY,
]

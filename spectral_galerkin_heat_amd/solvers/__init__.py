"""Solver selection (sgh/solvers/__init__.py:13-34) with the ``"mi355x"`` case."""

from .base import HeatSolver


def build_solver(context):
    """Return a solver for ``context.backend``.

    ``"mi355x"`` (also accepted as ``"gpu"``) gives ``SpectralSolver`` on the MI355X
    backend. The reference's ``"cpu"``, and ``"cpu_linear"`` belong to the reference
    package and raise here, as does any unknown name.
    """
    from .spectral import SpectralSolver
    match context.backend:
        case "mi355x" | "gpu":
            from ..backends import get_backend
            return SpectralSolver(backend=get_backend("mi355x"))
        case _:
            raise ValueError(
                f"Unknown backend: {context.backend!r}. This package provides 'mi355x' "
                "(alias 'gpu').")


__all__ = ["HeatSolver", "build_solver"]

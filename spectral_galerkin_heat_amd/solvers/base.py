"""Abstract heat-solver interface (sgh/solvers/base.py:31-91)."""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Any


class HeatSolver(ABC):
    """Call ``initialize`` once, ``step`` repeatedly, ``finalize`` at the end."""

    @abstractmethod
    def initialize(self, context) -> Any:
        ...

    @abstractmethod
    def step(self, t: float, dt: float) -> tuple:
        ...

    @abstractmethod
    def finalize(self) -> None:
        ...

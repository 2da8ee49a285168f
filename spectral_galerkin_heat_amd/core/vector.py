"""Immutable ``(x, y, z)`` triple used for domain sizes, mode counts and spacings.

Mirrors the reference's ``Vec3`` (``src/spectral_galerkin_heat/core/vector.py:36-110``):
iterable, indexable ``0 -> x``, element-wise ``/`` and ``*``, and ``zyx()`` for the
``[z, y, x]`` array-index order every solver array uses.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Iterator, Union

Number = Union[int, float]


@dataclass(frozen=True)
class Vec3:
    x: Number
    y: Number
    z: Number

    def __iter__(self) -> Iterator[Number]:
        return iter((self.x, self.y, self.z))

    def __getitem__(self, i: int) -> Number:
        return (self.x, self.y, self.z)[i]

    def __len__(self) -> int:
        return 3

    def zyx(self) -> "Vec3":
        return Vec3(self.z, self.y, self.x)

    def map(self, f: Callable[[Number], Number]) -> "Vec3":
        return Vec3(f(self.x), f(self.y), f(self.z))

    def _binary(self, other, op) -> "Vec3":
        if isinstance(other, Vec3):
            return Vec3(op(self.x, other.x), op(self.y, other.y), op(self.z, other.z))
        return Vec3(op(self.x, other), op(self.y, other), op(self.z, other))

    def __truediv__(self, other) -> "Vec3":
        return self._binary(other, lambda a, b: a / b)

    def __mul__(self, other) -> "Vec3":
        return self._binary(other, lambda a, b: a * b)

    def prod(self) -> Number:
        return self.x * self.y * self.z

"""Small SI unit table for ``{value, unit}`` config entries.

The reference converts config quantities with pint
(``src/spectral_galerkin_heat/core/parameters.py:44-121``): a value written as
``{value: 50, unit: "um"}`` is converted to the target unit (``"m"``), and an
incompatible unit raises. pint is not a dependency here, so this module parses the
handful of unit grammars the configs use -- ``*`` / ``.`` / ``/`` / ``^`` / ``**``,
parentheses, and the compact exponent form ``m2`` -- into a scale factor and a
dimension vector over ``(m, kg, s, K, mol)``.
"""

from __future__ import annotations

import re

__all__ = ["DimensionalityError", "UnitError", "convert", "parse_unit"]


class UnitError(ValueError):
    """Unknown unit name or malformed unit expression."""


class DimensionalityError(UnitError):
    """Source and target units measure different physical quantities."""


# name -> (scale to SI, (m, kg, s, K, mol))
_BASE = {
    "m": (1.0, (1, 0, 0, 0, 0)),
    "g": (1e-3, (0, 1, 0, 0, 0)),
    "s": (1.0, (0, 0, 1, 0, 0)),
    "K": (1.0, (0, 0, 0, 1, 0)),
    "mol": (1.0, (0, 0, 0, 0, 1)),
    "W": (1.0, (2, 1, -3, 0, 0)),
    "J": (1.0, (2, 1, -2, 0, 0)),
    "N": (1.0, (1, 1, -2, 0, 0)),
    "Pa": (1.0, (-1, 1, -2, 0, 0)),
    "Hz": (1.0, (0, 0, -1, 0, 0)),
    "bar": (1e5, (-1, 1, -2, 0, 0)),
    "atm": (101325.0, (-1, 1, -2, 0, 0)),
    "min": (60.0, (0, 0, 1, 0, 0)),
    "h": (3600.0, (0, 0, 1, 0, 0)),
    "L": (1e-3, (3, 0, 0, 0, 0)),
    "inch": (0.0254, (1, 0, 0, 0, 0)),
    "kelvin": (1.0, (0, 0, 0, 1, 0)),
    "meter": (1.0, (1, 0, 0, 0, 0)),
    "second": (1.0, (0, 0, 1, 0, 0)),
    "watt": (1.0, (2, 1, -3, 0, 0)),
    "joule": (1.0, (2, 1, -2, 0, 0)),
    "pascal": (1.0, (-1, 1, -2, 0, 0)),
}
_PREFIX = {
    "G": 1e9, "M": 1e6, "k": 1e3, "h": 1e2, "c": 1e-2, "m": 1e-3,
    "u": 1e-6, "µ": 1e-6, "μ": 1e-6, "n": 1e-9, "p": 1e-12,
}
_DIMENSIONLESS = (1.0, (0, 0, 0, 0, 0))

_TOKEN = re.compile(
    r"\s*(?:(?P<num>[-+]?\d+(?:\.\d+)?)|(?P<name>[A-Za-zµμ]+)(?P<exp>\d+)?"
    r"|(?P<op>\*\*|\^|\*|/|\.|\(|\)))"
)


def _lookup(name: str):
    if name in _BASE:
        return _BASE[name]
    if len(name) > 1 and name[0] in _PREFIX and name[1:] in _BASE:
        scale, dims = _BASE[name[1:]]
        return _PREFIX[name[0]] * scale, dims
    raise UnitError(f"Unknown unit {name!r}.")


def _pow(u, p):
    return u[0] ** p, tuple(d * p for d in u[1])


def _mul(a, b, sign=1):
    return (a[0] * (b[0] if sign > 0 else 1.0 / b[0]),
            tuple(x + sign * y for x, y in zip(a[1], b[1])))


class _Parser:
    def __init__(self, text: str):
        self.toks = []
        pos = 0
        text = text.strip()
        while pos < len(text):
            m = _TOKEN.match(text, pos)
            if m is None or m.end() == pos:
                raise UnitError(f"Cannot parse unit expression {text!r}.")
            pos = m.end()
            if m.group("num") is not None:
                self.toks.append(("num", float(m.group("num"))))
            elif m.group("name") is not None:
                self.toks.append(("name", m.group("name")))
                if m.group("exp") is not None:
                    self.toks.append(("op", "^"))
                    self.toks.append(("num", float(m.group("exp"))))
            else:
                self.toks.append(("op", m.group("op")))
        self.i = 0
        self.text = text

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def take(self):
        tok = self.peek()
        self.i += 1
        return tok

    def expr(self):
        u = self.factor()
        while True:
            kind, val = self.peek()
            if kind == "op" and val in ("*", "."):
                self.take()
                u = _mul(u, self.factor(), +1)
            elif kind == "op" and val == "/":
                self.take()
                u = _mul(u, self.factor(), -1)
            else:
                return u

    def factor(self):
        kind, val = self.take()
        if kind == "name":
            u = _lookup(val)
        elif kind == "num":
            u = (val, (0, 0, 0, 0, 0))
        elif kind == "op" and val == "(":
            u = self.expr()
            if self.take() != ("op", ")"):
                raise UnitError(f"Unbalanced parentheses in {self.text!r}.")
        else:
            raise UnitError(f"Cannot parse unit expression {self.text!r}.")
        kind, val = self.peek()
        if kind == "op" and val in ("^", "**"):
            self.take()
            k2, p = self.take()
            if k2 != "num":
                raise UnitError(f"Bad exponent in {self.text!r}.")
            u = _pow(u, p)
        return u


def parse_unit(text: str):
    """Return ``(scale_to_SI, dims)`` for a unit expression; ``""`` is dimensionless."""
    if text is None or str(text).strip() == "":
        return _DIMENSIONLESS
    p = _Parser(str(text))
    u = p.expr()
    if p.i != len(p.toks):
        raise UnitError(f"Trailing tokens in unit expression {text!r}.")
    return u


def convert(value: float, unit: str, target: str) -> float:
    """Convert ``value`` in ``unit`` to ``target``; raise on a dimension mismatch."""
    s_scale, s_dims = parse_unit(unit)
    t_scale, t_dims = parse_unit(target)
    if any(abs(a - b) > 1e-12 for a, b in zip(s_dims, t_dims)):
        raise DimensionalityError(
            f"Cannot convert from {unit!r} to {target!r}: incompatible dimensions.")
    return float(value) * s_scale / t_scale

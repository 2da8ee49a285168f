"""Temperature-dependent material properties (solid/liquid branches blended by f_l).

Mirrors ``src/spectral_galerkin_heat/core/properties.py``:
``liquid_fraction`` (:61-76), ``TempProperty`` (:79-140) and ``MaterialModel``
(:242-479). Host-side evaluation is kept for configuration checks and reference
constants; on the device the six branch expressions run as one packed stack program
(``MaterialModel.device_program``) inside the fused property kernels.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np

from .property_expr import MAX_CONSTS, MAX_OPS, PropertyExpr

__all__ = ["MaterialModel", "TempProperty", "liquid_fraction", "pack_property_program"]


def liquid_fraction(T, T_solidus, T_liquidus):
    """``f_l(T)`` clamped to ``[0, 1]``; a step at ``T_solidus`` if the band is empty.

    Reference: ``properties.py:61-76``.
    """
    if T_liquidus <= T_solidus:
        if hasattr(T, "shape"):
            return np.where(np.asarray(T) >= T_solidus, 1.0, 0.0).astype(T.dtype)
        return 1.0 if T >= T_solidus else 0.0
    f = (T - T_solidus) / (T_liquidus - T_solidus)
    if hasattr(f, "clip"):
        return f.clip(0.0, 1.0)
    return min(max(f, 0.0), 1.0)


@dataclass
class TempProperty:
    """One property with solid/liquid branch expressions (``properties.py:79-140``)."""

    solid: PropertyExpr
    liquid: PropertyExpr
    T_solidus: float
    T_liquidus: float
    is_constant: bool = False

    def __call__(self, T):
        if self.is_constant:
            return self.solid(T)
        ps = self.solid(T)
        pl = self.liquid(T)
        fl = liquid_fraction(T, self.T_solidus, self.T_liquidus)
        return (1.0 - fl) * ps + fl * pl

    @classmethod
    def from_config(cls, spec: Any, T_solidus: float, T_liquidus: float) -> "TempProperty":
        if isinstance(spec, dict) and ("solid" in spec or "liquid" in spec):
            solid = PropertyExpr(spec.get("solid", spec.get("liquid")))
            liquid = PropertyExpr(spec.get("liquid", spec.get("solid")))
            is_const = (solid.is_constant and liquid.is_constant
                        and solid.constant_value == liquid.constant_value)
            return cls(solid=solid, liquid=liquid, T_solidus=float(T_solidus),
                       T_liquidus=float(T_liquidus), is_constant=is_const)
        value = spec["value"] if isinstance(spec, dict) else spec
        const = PropertyExpr(value)
        return cls(solid=const, liquid=const, T_solidus=float(T_solidus),
                   T_liquidus=float(T_liquidus), is_constant=True)


# Layout of the packed device program (must match SghPropProgram in sgh_props.h):
#   int32 len[6]; int32 ops[6][MAX_OPS]; double consts[6][MAX_CONSTS];
#   double t_sol; double inv_band; int32 degenerate; int32 same
#   (bit b of `same`: the solid and liquid programs of property b — k, rho, c — are
#   identical, so the device evaluates one and blends it with itself)
# Branch order: k_solid, k_liquid, rho_solid, rho_liquid, c_solid, c_liquid.
PROGRAM_BYTES = 6 * 4 + 6 * MAX_OPS * 4 + 6 * MAX_CONSTS * 8 + 8 + 8 + 4 + 4


def pack_property_program(exprs, t_sol: float, t_liq: float) -> bytes:
    """Serialise six branch expressions plus the mushy band into the device layout."""
    assert len(exprs) == 6
    lens, ops_all, consts_all = [], [], []
    for e in exprs:
        ops, consts = e.to_program()
        lens.append(len(ops))
        ops_all.append(ops + [0] * (MAX_OPS - len(ops)))
        consts_all.append(consts + [0.0] * (MAX_CONSTS - len(consts)))
    degenerate = t_liq <= t_sol
    inv_band = 0.0 if degenerate else 1.0 / (t_liq - t_sol)
    same = 0
    for b in range(3):
        s_, l_ = 2 * b, 2 * b + 1
        if (lens[s_] == lens[l_] and ops_all[s_] == ops_all[l_]
                and struct.pack(f"<{MAX_CONSTS}d", *consts_all[s_]) ==
                struct.pack(f"<{MAX_CONSTS}d", *consts_all[l_])):
            same |= 1 << b
    blob = struct.pack("<6i", *lens)
    for ops in ops_all:
        blob += struct.pack(f"<{MAX_OPS}i", *ops)
    for consts in consts_all:
        blob += struct.pack(f"<{MAX_CONSTS}d", *consts)
    blob += struct.pack("<ddii", float(t_sol), float(inv_band), int(degenerate), same)
    assert len(blob) == PROGRAM_BYTES
    return blob


@dataclass
class MaterialModel:
    """``k(T)``, ``rho(T)``, ``c(T)`` and ``a(T) = rho·c`` (``properties.py:242-479``)."""

    k: TempProperty
    rho: TempProperty
    c: TempProperty

    def a(self, T):
        return self.rho(T) * self.c(T)

    @property
    def is_constant(self) -> bool:
        return self.k.is_constant and self.rho.is_constant and self.c.is_constant

    def band(self):
        """The single mushy band shared by k, rho and c (``properties.py:268-289``)."""
        bands = {(p.T_solidus, p.T_liquidus) for p in (self.k, self.rho, self.c)}
        if len(bands) != 1:
            raise ValueError(
                "k, rho and c must share one mushy band for the fused property "
                f"kernel; got bands {bands}.")
        return self.k.T_solidus, self.k.T_liquidus

    def device_program(self) -> bytes:
        t_sol, t_liq = self.band()
        return pack_property_program(
            [self.k.solid, self.k.liquid, self.rho.solid, self.rho.liquid,
             self.c.solid, self.c.liquid], t_sol, t_liq)

    def k_prime_a_prime(self, T, k_bar, a_bar):
        """Host form of the fluctuations ``(k(T) - k̄, a(T) - ā)``."""
        return self.k(T) - k_bar, self.a(T) - a_bar

    def recommended_references(self, T_lo: float, T_hi: float, n_samples: int = 512):
        """½(min + max) of each property over ``[T_lo, T_hi]`` (``properties.py:440-461``)."""
        T = np.linspace(float(T_lo), float(T_hi), int(n_samples))

        def _mid(prop):
            vals = np.asarray(prop(T), dtype=np.float64)
            return 0.5 * (float(vals.min()) + float(vals.max()))

        return _mid(self.k), _mid(self.rho), _mid(self.c)

    @classmethod
    def from_config(cls, mat_cfg: dict, T_solidus: float, T_liquidus: float
                    ) -> Optional["MaterialModel"]:
        def _has_branches(spec):
            return isinstance(spec, dict) and ("solid" in spec or "liquid" in spec)

        specs = {name: mat_cfg.get(key)
                 for name, key in (("k", "k"), ("rho", "rho"), ("c", "Cp"))}
        if not any(_has_branches(s) for s in specs.values()):
            return None
        return cls(
            k=TempProperty.from_config(specs["k"], T_solidus, T_liquidus),
            rho=TempProperty.from_config(specs["rho"], T_solidus, T_liquidus),
            c=TempProperty.from_config(specs["c"], T_solidus, T_liquidus),
        )

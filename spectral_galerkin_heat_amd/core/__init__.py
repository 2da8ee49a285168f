"""Configuration, laser and material-property layer (host side)."""

from .laser import (ConstantVelocityLaser, FixedLaser, LaserPath, LaserProfile, LaserState,
                    RasterLaserPath, SingleTrackLaser, build_laser_profile)
from .parameters import (FineMeshParams, GeomParams, LaserParams, MaterialParams, NumParams,
                         SimulationContext)
from .properties import MaterialModel, TempProperty, liquid_fraction
from .property_expr import PropertyExpr
from .vector import Vec3

__all__ = [
    "ConstantVelocityLaser", "FineMeshParams", "FixedLaser", "GeomParams", "LaserParams",
    "LaserPath", "LaserProfile", "LaserState", "MaterialModel", "MaterialParams", "NumParams",
    "PropertyExpr", "RasterLaserPath", "SimulationContext", "SingleTrackLaser", "TempProperty",
    "Vec3", "build_laser_profile", "liquid_fraction",
]

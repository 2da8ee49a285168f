"""Configuration dataclasses and the YAML-dict parser ``SimulationContext.from_dict``.

Same schema and semantics as ``src/spectral_galerkin_heat/core/parameters.py``:
``NumParams`` (:156-209), ``MaterialParams`` (:211-307), ``GeomParams`` (:309-331),
``LaserParams`` (:333-388), ``FineMeshParams`` (:390-410), the strict section/key
schema (:417-472) and ``from_dict`` (:513-682) -- including the precision casts of
material constants to the configured dtype, the ``n_steps = round(t_end/dt)`` step
rounding with ``dt = t_end/n_steps``, and the mandatory ``reference:`` key of a
temperature-dependent property. Units go through the small table in ``units.py``.
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from .properties import MaterialModel
from .units import convert
from .vector import Vec3

__all__ = [
    "FineMeshParams", "GeomParams", "LaserParams", "MaterialParams", "NumParams",
    "SimulationContext",
]


def _get_value(raw_val: Any, target_unit: Optional[str] = None) -> float:
    if isinstance(raw_val, dict):
        value = raw_val["value"]
        unit = raw_val.get("unit")
        if unit and target_unit:
            return convert(float(value), unit, target_unit)
        return float(value)
    if isinstance(raw_val, (int, float, np.floating, np.integer)) and not isinstance(raw_val, bool):
        return float(raw_val)
    raise TypeError(f"Cannot extract a value from type {type(raw_val)}")


def _get_vector(raw_val: Any, target_unit: Optional[str] = None) -> list:
    if isinstance(raw_val, dict):
        values = raw_val["value"]
        unit = raw_val.get("unit")
        if unit and target_unit:
            return [convert(float(v), unit, target_unit) for v in values]
        return [float(v) for v in values]
    if isinstance(raw_val, (list, tuple)):
        return [float(_get_value(v, target_unit)) for v in raw_val]
    raise TypeError(f"Cannot extract a vector from type {type(raw_val)}")


_SUPPORTED_DTYPES = {"float32": np.float32, "float64": np.float64}


def _resolve_dtype(name):
    key = str(name).lower()
    try:
        return _SUPPORTED_DTYPES[key]
    except KeyError:
        choices = ", ".join(sorted(_SUPPORTED_DTYPES))
        raise ValueError(f"Unknown dtype: {name!r}. Choose one of: {choices}.") from None


@dataclass
class NumParams:
    dt: float
    nx: int
    ny: int
    nz: int
    t_end: float = 0.0
    n_steps: int = 0
    dt_nominal: float = 0.0
    update_interval: float = 1e-3
    save_all: bool = False
    max_picard_iter: Optional[int] = None
    picard_tol: Optional[float] = None
    picard_omega: Optional[float] = None
    dtype: Any = np.float32
    # arithmetic of the 3-D cosine transforms on the MI355X backend: "mfma" (matrix cores
    # in dtype), "bf16x3" (split-BF16 products, float32 only) or "valu" (vector ALU)
    transform: str = "mfma"


@dataclass
class MaterialParams:
    name: str = "Material"
    rho: float = 1.0
    k: float = 1.0
    Cp: float = 1.0
    L_f: float = 0.0
    T_solidus: float = 0.0
    T_liquidus: float = 0.0
    Pa: float = 0
    R_v: float = 0
    T_boil: float = 0
    DeltaH_LV: float = 0
    T0: float = 0
    h_conv_top: float = 0.0
    h_conv_bottom: float = 0.0
    model: Optional[MaterialModel] = None

    @property
    def diff(self) -> float:
        return self.k / (self.rho * self.Cp)


@dataclass
class GeomParams:
    size: Vec3
    n: Vec3
    d: Vec3 = field(init=False)

    def __post_init__(self):
        self.d = self.size / self.n


@dataclass
class LaserParams:
    radius: float
    absorptivity: float
    power: float = 0.0
    profile: str = "gaussian"
    r_x: float = 0.0
    r_y: float = 0.0
    super_gaussian_order: float = 2.0
    cell_integrated: bool = False

    def __post_init__(self):
        if not self.r_x:
            self.r_x = self.radius
        if not self.r_y:
            self.r_y = self.radius

    @property
    def ref_area(self) -> float:
        return self.r_x * self.r_y


@dataclass
class FineMeshParams:
    refinement: int = 4
    box_size: Vec3 = Vec3(0.9e-3, 0.9e-3, 0.04e-3)


_SECTION_KEYS = {
    "simulation": {
        "name", "backend", "dtype", "dt", "duration", "update_interval",
        "max_picard_iter", "picard_tol", "picard_omega", "transform",
    },
    "domain": {"size", "mesh"},
    "material": {
        "name", "rho", "k", "Cp", "L_f", "T_solidus", "T_liquidus", "T0",
        "Pa", "R_v", "T_boil", "DeltaH_LV", "h_conv_top", "h_conv_bottom",
    },
    "laser": {
        "radius", "absorptivity", "power_nominal", "profile", "r_x", "r_y",
        "super_gaussian_order", "cell_integrated", "path",
    },
    "fine_mesh": {"refinement", "box_size"},
}
_PASSTHROUGH_SECTIONS = {"io", "post_processing"}
_LASER_PATH_KEYS = {"type", "file", "initial_position"}


def _reject_unknown(section: str, keys, allowed) -> None:
    for key in keys:
        if key not in allowed:
            raise ValueError(
                f"Unknown key {key!r} in the {section!r} config section. "
                f"Valid keys: {', '.join(sorted(allowed))}.")


def _validate_config(cfg: dict) -> None:
    known = set(_SECTION_KEYS) | _PASSTHROUGH_SECTIONS
    for section in cfg:
        if section not in known:
            raise ValueError(
                f"Unknown config section {section!r}. "
                f"Valid sections: {', '.join(sorted(known))}.")
    for section, allowed in _SECTION_KEYS.items():
        body = cfg.get(section)
        if isinstance(body, dict):
            _reject_unknown(section, body, allowed)
    laser = cfg.get("laser")
    path_cfg = laser.get("path") if isinstance(laser, dict) else None
    if isinstance(path_cfg, dict):
        _reject_unknown("laser.path", path_cfg, _LASER_PATH_KEYS)


@dataclass
class SimulationContext:
    num: NumParams
    mat: MaterialParams
    geom: GeomParams
    laser: LaserParams
    laser_path: Any
    io: dict
    backend: str = "cpu"
    fine: Optional[FineMeshParams] = field(default_factory=FineMeshParams)

    @classmethod
    def from_dict(cls, cfg: dict, config_dir: Optional[str] = None) -> "SimulationContext":
        _validate_config(cfg)
        real_t = _resolve_dtype(cfg.get("simulation", {}).get("dtype", "float32"))
        sim_cfg = cfg.get("simulation", {})
        domain_cfg = cfg.get("domain", {})
        sim_backend = str(sim_cfg.get("backend", "cpu")).lower()

        Lx, Ly, Lz = _get_vector(domain_cfg["size"], "m")
        nx, ny, nz = _get_vector(domain_cfg["mesh"])

        t_end = float(_get_value(sim_cfg.get("duration", 0.01), "s"))
        dt_nominal = float(real_t(_get_value(sim_cfg["dt"], "s")))
        n_steps = round(t_end / dt_nominal)
        dt = t_end / n_steps

        num = NumParams(
            dt=dt, nx=int(nx), ny=int(ny), nz=int(nz), t_end=t_end, n_steps=n_steps,
            dt_nominal=dt_nominal,
            update_interval=float(_get_value(sim_cfg.get("update_interval", 1e-3), "s")),
            max_picard_iter=(int(sim_cfg["max_picard_iter"])
                             if sim_cfg.get("max_picard_iter") is not None else None),
            picard_tol=(float(sim_cfg["picard_tol"])
                        if sim_cfg.get("picard_tol") is not None else None),
            picard_omega=(float(sim_cfg["picard_omega"])
                          if sim_cfg.get("picard_omega") is not None else None),
            dtype=real_t,
            transform=str(sim_cfg.get("transform", "mfma")).lower(),
        )
        geom = GeomParams(size=Vec3(Lx, Ly, Lz), n=Vec3(int(nx), int(ny), int(nz)))

        fine_cfg = cfg.get("fine_mesh", None)
        if fine_cfg is None:
            fine = None
        else:
            box_val = fine_cfg.get("box_size", (0.9e-3, 0.9e-3, 0.04e-3))
            box = [float(_get_value(v, "m")) for v in box_val]
            fine = FineMeshParams(refinement=int(_get_value(fine_cfg.get("refinement", 4))),
                                  box_size=Vec3(*box))

        mat_cfg = cfg.get("material", {})
        T_solidus = real_t(_get_value(mat_cfg.get("T_solidus", 0.0), "K"))
        T_liquidus = real_t(_get_value(mat_cfg.get("T_liquidus", 0.0), "K"))
        T0 = real_t(_get_value(mat_cfg.get("T0", 0.0), "K"))
        T_boil = real_t(_get_value(mat_cfg.get("T_boil", 0.0), "K"))

        model = MaterialModel.from_config(mat_cfg, float(T_solidus), float(T_liquidus))
        if model is None:
            rho_ref = real_t(_get_value(mat_cfg["rho"], "kg / m**3"))
            k_ref = real_t(_get_value(mat_cfg["k"], "W / (m * K)"))
            cp_ref = real_t(_get_value(mat_cfg["Cp"], "J / (kg * K)"))
        else:
            refs, missing = {}, []
            for key, tprop, unit in (("k", model.k, "W / (m * K)"),
                                     ("rho", model.rho, "kg / m**3"),
                                     ("Cp", model.c, "J / (kg * K)")):
                spec = mat_cfg.get(key)
                is_branch = isinstance(spec, dict) and ("solid" in spec or "liquid" in spec)
                if is_branch and not tprop.is_constant:
                    if "reference" not in spec:
                        missing.append(key)
                        continue
                    refs[key] = real_t(_get_value(spec["reference"], unit))
                else:
                    refs[key] = real_t(float(tprop(float(T0))))
            if missing:
                rec = dict(zip(("k", "rho", "Cp"),
                               model.recommended_references(float(T0), float(T_boil))))
                rec_str = ", ".join(f"{k}.reference={rec[k]:g}" for k in missing)
                raise ValueError(
                    f"Temperature-dependent material property {missing} needs an "
                    f"explicit 'reference:' key inside its block. The recommended "
                    f"values, each the average of that property's two extrema over "
                    f"[T0, T_boil] K, are: {rec_str}.")
            k_ref, rho_ref, cp_ref = refs["k"], refs["rho"], refs["Cp"]

        mat = MaterialParams(
            name=mat_cfg.get("name", "Material"), rho=rho_ref, k=k_ref, Cp=cp_ref,
            L_f=real_t(_get_value(mat_cfg.get("L_f", 0.0), "J / kg")),
            T_solidus=T_solidus, T_liquidus=T_liquidus,
            Pa=real_t(_get_value(mat_cfg.get("Pa", 0.0), "Pa")),
            R_v=real_t(_get_value(mat_cfg.get("R_v", 0.0), "J / (kg * K)")),
            T_boil=T_boil,
            DeltaH_LV=real_t(_get_value(mat_cfg.get("DeltaH_LV", 0.0), "J / kg")),
            T0=T0,
            h_conv_top=real_t(_get_value(mat_cfg.get("h_conv_top", 0.0), "W / (m**2 * K)")),
            h_conv_bottom=real_t(_get_value(mat_cfg.get("h_conv_bottom", 0.0), "W / (m**2 * K)")),
            model=model,
        )

        laser_cfg = cfg.get("laser", {})
        laser = LaserParams(
            radius=real_t(_get_value(laser_cfg["radius"], "m")),
            absorptivity=real_t(_get_value(laser_cfg["absorptivity"])),
            power=real_t(_get_value(laser_cfg["power_nominal"], "W")),
            profile=str(laser_cfg.get("profile", "gaussian")),
            r_x=real_t(_get_value(laser_cfg.get("r_x", 0.0), "m")),
            r_y=real_t(_get_value(laser_cfg.get("r_y", 0.0), "m")),
            super_gaussian_order=float(_get_value(laser_cfg.get("super_gaussian_order", 2.0))),
            cell_integrated=bool(laser_cfg.get("cell_integrated", False)),
        )

        path_cfg = laser_cfg.get("path", {}) or {}
        laser_path = None
        if str(path_cfg.get("type", "")).lower() == "gcode":
            from ..io_utils.gcode_path import GCodeLaserPath
            gcode_file = path_cfg.get("file", None)
            initial_position = tuple(path_cfg.get("initial_position", [0.0, 0.0]))
            if gcode_file is not None:
                if not os.path.isabs(gcode_file):
                    if config_dir is not None:
                        gcode_file = os.path.abspath(os.path.join(config_dir, "paths", gcode_file))
                    else:
                        gcode_file = os.path.abspath(os.path.join(
                            os.getcwd(), "simulations", "examples", "paths", gcode_file))
                cfg["laser"]["path"]["file"] = gcode_file
                laser_path = GCodeLaserPath(gcode_file, initial_position=initial_position)

        return cls(num=num, mat=mat, geom=geom, laser=laser, laser_path=laser_path,
                   io=cfg.get("io", {}), backend=sim_backend, fine=fine)

"""Scan-simulation workloads: the single-track, raster and batched cases.

Each builder returns a reference-schema config dict (the same YAML layout
``SimulationContext.from_dict`` reads) and the ``LaserPath`` for the scan. Materials
are 316L-like; evaporation uses the reference's law (prefactor 0.82, gated at the
liquidus; ``physics/spectral_cpu_kernels.py:144-162``) with ``R_v = R/M``.
"""

from __future__ import annotations

import copy

import numpy as np

from .core.laser import RasterLaserPath, SingleTrackLaser
from .core.parameters import SimulationContext

__all__ = [
    "MATERIAL_316L", "MATERIAL_316L_LATENT", "batch_uneven_cuboids", "make_context",
    "single_track_box", "single_track_constant", "single_track_tdep", "tdep_track_box",
    "raster_multitrack", "serving_request", "serving_stream",
]

R_GAS = 8.314462618
M_316L = 0.0558

MATERIAL_316L = {
    "name": "316L", "rho": 7900.0, "k": 20.0, "Cp": 500.0, "T0": 293.0,
    "L_f": 0.0, "T_solidus": 0.0, "T_liquidus": 0.0,
    "Pa": 101325.0, "T_boil": 3090.0, "DeltaH_LV": 7.45e6, "R_v": R_GAS / M_316L,
    "h_conv_top": 0.0, "h_conv_bottom": 0.0,
}


# Constant 316L at T0 = 293 K with latent heat, bottom convection and the evaporation
# constants of the reference's example configurations (02_single_track.yaml).
MATERIAL_316L_LATENT = {
    "name": "316L", "rho": 7957.5, "k": 13.851, "Cp": 497.89, "T0": 293.0,
    "L_f": 2.677e5, "T_solidus": 1674.15, "T_liquidus": 1697.15,
    "Pa": 101325.0, "T_boil": 3090.0, "DeltaH_LV": 7.416e6, "R_v": 150.774,
    "h_conv_top": 0.0, "h_conv_bottom": 3000.0,
}


def make_context(cfg: dict, laser_path) -> SimulationContext:
    ctx = SimulationContext.from_dict(copy.deepcopy(cfg))
    ctx.laser_path = laser_path
    return ctx


def _cfg(size, mesh, dt, n_steps, mat, radius, absorptivity, power, tol, maxit,
         dtype="float64", omega=None, transform=None):
    sim = {"backend": "mi355x", "dtype": dtype, "dt": dt, "duration": dt * n_steps,
           "picard_tol": tol, "max_picard_iter": maxit}
    if omega is not None:
        sim["picard_omega"] = omega
    if transform is not None:
        sim["transform"] = transform
    return {
        "simulation": sim,
        "domain": {"size": list(size), "mesh": list(mesh)},
        "material": copy.deepcopy(mat),
        "laser": {"radius": radius, "absorptivity": absorptivity, "power_nominal": power},
        "io": {},
    }


def single_track_constant(mesh=(64, 32, 16), n_steps=320, dtype="float64"):
    """Straight track on a 2.0×1.0×0.5 mm cuboid, constant properties, evaporation only.

    Gaussian P = 200 W, 1/e² radius 50 µm, A = 0.35, 0.5 m/s along x at mid-y from
    x = 0.2 mm; Δt = 1e-5 s; Picard tol 1e-8, cap 20, ω = 0.1 (the solver default).
    """
    cfg = _cfg((2.0e-3, 1.0e-3, 0.5e-3), mesh, 1e-5, n_steps, MATERIAL_316L,
               50e-6, 0.35, 200.0, 1e-8, 20, dtype)
    return cfg, SingleTrackLaser(0.2e-3, 1.8e-3, 0.5e-3, 0.5, 200.0)


def single_track_tdep(mesh=(128, 64, 32), n_steps=640, dtype="float64", transform=None):
    """The same track with latent heat, top/bottom convection and T-dependent k, cp.

    L_f = 2.7e5 J/kg over 1658–1723 K (grid-mode latent source), h = 10 W/m²K to
    293 K on z = 0 and z = Lz, k(T) = min(14.6 + 0.0127(T − 293), 30),
    cp(T) = min(450 + 0.28(T − 293), 750) with the caps written through ``fabs``,
    reference k̄ = 22.3, c̄ = 600 (midpoints of their ranges); Δt = 5e-6 s, tol 1e-6.
    ``transform`` selects the 3-D transform arithmetic ("mfma", "bf16x3", "valu").
    """
    mat = copy.deepcopy(MATERIAL_316L)
    k_lin, cp_lin = "(14.6 + 0.0127*(T - 293.0))", "(450.0 + 0.28*(T - 293.0))"
    k_expr = f"0.5*({k_lin} + 30.0 - fabs({k_lin} - 30.0))"
    cp_expr = f"0.5*({cp_lin} + 750.0 - fabs({cp_lin} - 750.0))"
    mat.update({
        "L_f": 2.7e5, "T_solidus": 1658.0, "T_liquidus": 1723.0,
        "h_conv_top": 10.0, "h_conv_bottom": 10.0,
        "k": {"solid": k_expr, "liquid": k_expr, "reference": 22.3},
        "Cp": {"solid": cp_expr, "liquid": cp_expr, "reference": 600.0},
    })
    cfg = _cfg((2.0e-3, 1.0e-3, 0.5e-3), mesh, 5e-6, n_steps, mat,
               50e-6, 0.35, 200.0, 1e-6, 20, dtype, transform=transform)
    return cfg, SingleTrackLaser(0.2e-3, 1.8e-3, 0.5e-3, 0.5, 200.0)


def raster_multitrack(mesh=(256, 128, 64), n_steps=5000, dtype="float64", n_tracks=10,
                      size=(4.0e-3, 2.0e-3, 1.0e-3), power=250.0):
    """10 alternating x-tracks (1.2 → 2.8 mm), hatch 80 µm, 0.5 ms dwell, P = 250 W,
    r = 45 µm, 0.8 m/s, on a 4×2×1 mm cuboid; Δt = 5e-6 s; tol 1e-8, cap 20."""
    cfg = _cfg(size, mesh, 5e-6, n_steps, MATERIAL_316L, 45e-6, 0.35, power, 1e-8, 20, dtype)
    x0, x1 = 0.3 * size[0], 0.7 * size[0]
    return cfg, RasterLaserPath(n_tracks, x0, x1, 0.5 * size[1], 80e-6, 0.8, power, 0.5e-3)


def batch_uneven_cuboids(n=64, seed=0, n_steps=1000, dtype="float64"):
    """64 independent cuboids with per-axis mode counts drawn from {32,48,64,96,128}
    (z from {16,24,32}), extents 0.5–4 mm (x, y) and 0.25–1 mm (z), P 100–400 W,
    radius 35–70 µm, speed 0.2–1.5 m/s, one track at a random y in the middle 60 %,
    x from 0.1·Lx to 0.9·Lx then off; Δt = 5e-6 s; tol 1e-8, cap 20.

    Returns a list of ``(cfg, laser_path)``.
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx, ny = (int(v) for v in rng.choice([32, 48, 64, 96, 128], size=2))
        nz = int(rng.choice([16, 24, 32]))
        Lx, Ly = (float(v) for v in rng.uniform(0.5e-3, 4.0e-3, size=2))
        Lz = float(rng.uniform(0.25e-3, 1.0e-3))
        P = float(rng.uniform(100.0, 400.0))
        r = float(rng.uniform(35e-6, 70e-6))
        v = float(rng.uniform(0.2, 1.5))
        y = float(rng.uniform(0.2, 0.8)) * Ly
        cfg = _cfg((Lx, Ly, Lz), (nx, ny, nz), 5e-6, n_steps, MATERIAL_316L, r, 0.35, P,
                   1e-8, 20, dtype)
        out.append((cfg, SingleTrackLaser(0.1 * Lx, 0.9 * Lx, y, v, P)))
    return out


def serving_request(rng, dtype="float32"):
    """One scan request: batched-cuboid shape, or with probability 0.1 a
    256×128×64 raster of 1–5 tracks; 200–2000 steps."""
    n_steps = int(rng.integers(200, 2001))
    if rng.random() < 0.1:
        n_tracks = int(rng.integers(1, 6))
        cfg, lp = raster_multitrack(n_steps=n_steps, dtype=dtype, n_tracks=n_tracks)
        return cfg, lp
    (cfg, lp), = batch_uneven_cuboids(n=1, seed=int(rng.integers(1 << 31)), n_steps=n_steps,
                                      dtype=dtype)
    return cfg, lp


def serving_stream(n, rate_hz, seed, dtype="float32"):
    """``n`` scan requests drawn by ``serving_request`` from ``default_rng(seed)``, arriving
    as a Poisson process of ``rate_hz`` (first at t = 0, gaps from ``default_rng(seed + 1)``).

    Returns a list of ``(cfg, laser_path, n_steps, arrival_s)``.
    """
    rng = np.random.default_rng(seed)
    shapes = [serving_request(rng, dtype=dtype) for _ in range(n)]
    gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate_hz, size=n)
    gaps[0] = 0.0
    arrivals = np.cumsum(gaps)
    out = []
    for (cfg, lp), t in zip(shapes, arrivals):
        sim = cfg["simulation"]
        steps = int(round(sim["duration"] / sim["dt"]))
        out.append((cfg, lp, steps, float(t)))
    return out


def single_track_box(mesh=(80, 48, 120), n_steps=140, dtype="float64", refinement=2,
                     box_size=(0.6e-3, 0.6e-3, 0.07e-3), size=(0.6e-3, 0.4e-3, 0.15e-3)):
    """Box-mode single track (the reference's 02_single_track.yaml case).

    0.6×0.4×0.15 mm cuboid, 80×48×120 modes, constant 316L with latent heat and bottom
    convection (h = 3000 W/m²K); 120 W, r = 60 µm, A = 0.3, 0.4 m/s along x at
    y = 0.2 mm from x = 0.15 to 0.5 mm; latent heat on a laser-following fine box
    (refinement 2, 0.6×0.6×0.07 mm, clipped to the domain); Δt = 5e-6 s, 140 steps,
    Picard ω 0.1, tol 1e-4, cap 30 (the solver defaults).
    """
    cfg = _cfg(size, mesh, 5e-6, n_steps, MATERIAL_316L_LATENT, 60e-6, 0.30, 120.0, 1e-4, 30,
               dtype)
    cfg["fine_mesh"] = {"refinement": int(refinement), "box_size": list(box_size)}
    y = 0.5 * size[1]
    return cfg, SingleTrackLaser(0.25 * size[0], min(0.5e-3, 0.85 * size[0]), y, 0.4, 120.0)


def tdep_track_box(mesh=(100, 60, 96), n_steps=50, dtype="float64", refinement=2,
                   box_size=(0.2e-3, 0.15e-3, 0.04e-3), size=(0.5e-3, 0.3e-3, 0.12e-3)):
    """Box mode with temperature-dependent properties (the reference's
    03_nonlinear_tdep.yaml case): k, ρ, cp solid/liquid expressions of T blended by the
    liquid fraction, latent heat on a fine box (refinement 2, 0.2×0.15×0.04 mm), bottom
    convection; 24 W, r = 30 µm, A = 0.3, 0.15 m/s from x = 0.2 mm at mid-y;
    Δt = 4e-6 s, Picard ω 0.31, tol 1e-4, cap 80.
    """
    mat = copy.deepcopy(MATERIAL_316L_LATENT)
    mat.update({
        "k": {"solid": "9.248 + 0.01571 * T", "liquid": "12.41 + 0.003279 * T",
              "reference": 24.5075},
        "rho": {"solid": "8084.2 - 0.42086 * T - 3.8942e-5 * T**2",
                "liquid": "7432.7 + 0.039338 * T - 1.8007e-4 * T**2",
                "reference": 6896.2365},
        "Cp": {"solid": "458.98 + 0.1328 * T", "liquid": "769.86", "reference": 633.8752},
    })
    cfg = _cfg(size, mesh, 4e-6, n_steps, mat, 30e-6, 0.30, 24.0, 1e-4, 80, dtype, omega=0.31)
    cfg["fine_mesh"] = {"refinement": int(refinement), "box_size": list(box_size)}
    return cfg, SingleTrackLaser(0.4 * size[0], 0.8 * size[0], 0.5 * size[1], 0.15, 24.0)

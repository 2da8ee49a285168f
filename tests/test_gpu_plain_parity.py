"""GPU parity of the plain (reference-structured) MI355X path against the FP64 oracle.

Kernel-level checks run every HIP entry point once on seeded inputs and compare with
the oracle function it replaces (FP64: relative 1e-12 unless stated). Step-level checks
advance SpectralSolver on the mi355x backend and the oracle side by side and require
relative L2 of the modal state ≤ 1e-9 (FP64) with identical Picard counts, or ≤ 1e-4
against the FP64 oracle for the FP32 form.
"""

import copy
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import sgh_oracle as o  # noqa: E402
from spectral_galerkin_heat_amd.core.laser import ConstantVelocityLaser, FixedLaser  # noqa: E402
from spectral_galerkin_heat_amd.workloads import (make_context, raster_multitrack,  # noqa: E402
                                                  single_track_constant, single_track_tdep)

DEV = "cuda"


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, dtype)


def host(t):
    return t.detach().cpu().double().numpy()


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def K():
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as k
    return k


def _state(K, mesh=(12, 10, 8), size=(1.0e-3, 0.7e-3, 0.4e-3), dtype=np.float64, **mat):
    from spectral_galerkin_heat_amd.core.parameters import GeomParams, MaterialParams, NumParams
    from spectral_galerkin_heat_amd.core.vector import Vec3
    geom = GeomParams(size=Vec3(*size), n=Vec3(*mesh))
    num = NumParams(dt=1e-6, nx=mesh[0], ny=mesh[1], nz=mesh[2], dtype=dtype)
    phys = MaterialParams(rho=7900.0, k=20.0, Cp=500.0, **mat)
    return K.SpectralSolverState(phys, geom, num, None), geom, num, phys


# --- kernel level ---------------------------------------------------------------------

def test_etd_tables_match_oracle(K):
    S, geom, num, phys = _state(K, mesh=(16, 12, 8))
    Kr, KKr = o.precompute_K_KK(20.0, 7900.0, 500.0, 1e-6, (8, 12, 16), (0.4e-3, 0.7e-3, 1.0e-3))
    assert rel(host(S.K), Kr) < 1e-14
    np.testing.assert_allclose(host(S.KK), KKr, rtol=1e-11)


@pytest.mark.parametrize("shape", [(6, 5), (32, 64), (48, 96), (128, 256)])
def test_dct2d_matches_scipy(K, shape):
    rng = np.random.default_rng(0)
    q = rng.standard_normal(shape)
    out = host(K.DCT_II(dev(q)))
    assert rel(out, o.dct2(q)) < 1e-13
    assert rel(host(K.IDCT_II(dev(q))), o.idct2(q)) < 1e-13


@pytest.mark.parametrize("shape", [(4, 5, 6), (16, 32, 64), (32, 48, 96)])
def test_dct3d_matches_scipy(K, shape):
    rng = np.random.default_rng(1)
    f = rng.standard_normal(shape)
    assert rel(host(K.DCT_II(dev(f))), o.dct2(f)) < 1e-13
    assert rel(host(K.IDCT_II(dev(f))), o.idct2(f)) < 1e-13


def test_surface_and_volume_reconstruction(K):
    S, geom, *_ = _state(K, mesh=(24, 20, 12), size=(1.2e-3, 0.9e-3, 0.5e-3))
    g = o.OracleGrid(tuple(geom.size), tuple(geom.n))
    rng = np.random.default_rng(2)
    a = rng.standard_normal((12, 20, 24))
    a[0, 0, 0] += 500.0
    assert rel(host(K.reconstruct_surface_temperature(dev(a), S)), o.reconstruct_surface(a, g)) < 1e-13
    assert rel(host(K.reconstruct_bottom_temperature(dev(a), S)), o.reconstruct_bottom(a, g)) < 1e-13
    assert rel(host(K.reconstruct_volume(dev(a), S)), o.reconstruct_volume(a, g)) < 1e-13
    f = rng.standard_normal((12, 20, 24))
    assert rel(host(K.project_volume(dev(f), S)), o.project_volume(f, g)) < 1e-13


def test_pointwise_updates_hand_values(K):
    aK = dev(np.ones((2, 2, 2)))
    KK = dev(np.ones((2, 2, 2)))
    out = torch.empty_like(aK)
    K.update_modes_etd1(aK, KK, dev(np.array([2.0, 3.0])).view(2, 1, 1), dev(np.full((2, 2), 5.0)), out)
    assert np.all(host(out)[0] == 11.0) and np.all(host(out)[1] == 16.0)
    a = dev(np.ones((2, 2, 2)))
    K.add_source_term_modes(a, dev(np.full((2, 2, 2), 2.0)), dev(np.full((2, 2, 2), 3.0)))
    assert np.all(host(a) == 7.0)
    a = dev(np.ones((2, 2, 2)))
    K.add_bottom_surface_source(a, KK, dev(np.array([2.0, 4.0])).view(2, 1, 1), dev(np.full((2, 2), 5.0)))
    assert np.all(host(a)[0] == 11.0) and np.all(host(a)[1] == 21.0)


def test_evaporation_and_shift(K):
    rng = np.random.default_rng(4)
    T = rng.uniform(1000.0, 4000.0, (10, 12))
    q = torch.empty((10, 12), dtype=torch.float64, device=DEV)
    K.compute_evaporation_flux(dev(T), q, 101325.0, 3090.0, 7.45e6, 149.0, 1800.0)
    np.testing.assert_allclose(host(q), o.evaporation_flux(T, 101325.0, 3090.0, 7.45e6, 149.0, 1800.0),
                               rtol=1e-13, atol=0)
    from spectral_galerkin_heat_amd.core.parameters import GeomParams
    from spectral_galerkin_heat_amd.core.vector import Vec3
    geom = GeomParams(size=Vec3(12e-4, 10e-4, 1e-4), n=Vec3(12, 10, 4))
    f = rng.standard_normal((10, 12))
    for sh in [(0.3e-4, -0.7e-4), (-1.6e-5, 0.0), (2.5e-4, 1.0e-4)]:
        got = host(K.shift_flux(dev(f), sh, geom))
        np.testing.assert_allclose(got, o.shift_flux(f, sh, 1e-4, 1e-4), atol=1e-14)


@pytest.mark.parametrize("profile,order,cell", [("gaussian", 2.0, False), ("gaussian", 2.0, True),
                                                ("super_gaussian", 6.0, False)])
def test_laser_flux(K, profile, order, cell):
    from spectral_galerkin_heat_amd.core.laser import build_laser_profile
    from spectral_galerkin_heat_amd.core.parameters import LaserParams
    S, geom, *_ = _state(K, mesh=(40, 24, 8), size=(2e-3, 1e-3, 0.5e-3))
    g = o.OracleGrid(tuple(geom.size), tuple(geom.n))
    prof = build_laser_profile(profile, order, cell_integrated=cell)
    lp = LaserParams(radius=50e-6, absorptivity=0.35, r_x=50e-6, r_y=70e-6)
    peak = prof.peak_intensity(200.0, 0.35, lp.ref_area)
    got = host(K.laser_flux(S, prof, lp, 0.77e-3, 0.41e-3, peak))
    if cell:
        ref = o.gaussian_cell_integrated_flux(g.x, g.y, 0.77e-3, 0.41e-3, 50e-6, 70e-6, g.dx, g.dy, peak)
    else:
        ref = o.super_gaussian_flux(g.x, g.y, 0.77e-3, 0.41e-3, 50e-6, 70e-6, order, peak)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * ref.max())


def test_latent_and_property_correction(K):
    cfg, _ = single_track_tdep(mesh=(20, 12, 10))
    ctx = make_context(cfg, FixedLaser(1e-3, 5e-4, 200.0))
    S = K.SpectralSolverState(ctx.mat, ctx.geom, ctx.num, None)
    g = o.OracleGrid(tuple(ctx.geom.size), tuple(ctx.geom.n))
    props = o.OracleProps.from_model(ctx.mat.model)
    rng = np.random.default_rng(7)
    decay = np.exp(-(np.arange(20) / 5.0) ** 2)[None, None, :] * np.exp(-(np.arange(12) / 5.0) ** 2)[None, :, None] \
        * np.exp(-(np.arange(10) / 5.0) ** 2)[:, None, None]
    a = rng.standard_normal((10, 12, 20)) * 40.0 * decay
    a[0, 0, 0] = 1500.0 * math.sqrt(2e-3 * 1e-3 * 0.5e-3)
    T = o.reconstruct_volume(a, g)
    Tp = T - rng.uniform(0.0, 80.0, T.shape)
    Q = torch.empty((10, 12, 20), dtype=torch.float64, device=DEV)
    K.compute_latent_heat_source_grid(Q, dev(T), dev(Tp), ctx.mat, ctx.num, S)
    np.testing.assert_allclose(host(Q), props.latent(T, Tp, 2.7e5, ctx.num.dt), rtol=1e-12,
                               atol=1e-9 * np.abs(props.latent(T, Tp, 2.7e5, ctx.num.dt)).max())
    C = K.assemble_property_correction(S, dev(a), dev(Tp), ctx.num.dt, ctx.mat.model,
                                       float(ctx.mat.k), float(ctx.mat.rho) * float(ctx.mat.Cp))
    Cr = o.assemble_property_correction(a, Tp, ctx.num.dt, props, float(ctx.mat.k),
                                        float(ctx.mat.rho) * float(ctx.mat.Cp), g)
    assert rel(host(C), Cr) < 1e-11


# --- step level -----------------------------------------------------------------------

def _solver(cfg, lp):
    from spectral_galerkin_heat_amd.backends import get_backend
    from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
    ctx = make_context(cfg, lp)
    s = SpectralSolver(get_backend("mi355x"), ctx)
    s.initialize()
    return s, ctx


def _run_pair(cfg, lp, n_steps, track=False):
    s, ctx = _solver(cfg, lp)
    ref = o.OracleSolver(make_context(cfg, lp)).initialize()
    dt = ctx.num.dt
    its, its_ref = [], []
    for n in range(n_steps):
        _, m = s.step(n * dt, dt)
        _, mr = ref.step(n * dt, dt)
        its.append(m["n_evap_iter"])
        its_ref.append(mr["n_evap_iter"])
    return s, ref, its, its_ref, m, mr


def test_single_track_fp64_matches_oracle_1e9():
    cfg, lp = single_track_constant(n_steps=40)
    s, ref, its, its_ref, m, mr = _run_pair(cfg, lp, 40)
    assert its == its_ref
    assert rel(host(s.state.a), ref.a) < 1e-9
    assert float(m["T_surface_max"]) == pytest.approx(mr["T_surface_max"], rel=1e-9)
    assert float(m["P_laser"]) == pytest.approx(mr["P_laser"], rel=1e-12)
    assert m["picard_rel_err"] == pytest.approx(mr["picard_rel_err"], rel=1e-6)


def test_tdep_latent_convection_fp64_matches_oracle_1e9():
    cfg, lp = single_track_tdep(mesh=(48, 24, 16), n_steps=12)
    s, ref, its, its_ref, m, mr = _run_pair(cfg, lp, 12)
    assert its == its_ref
    assert rel(host(s.state.a), ref.a) < 1e-9
    assert rel(host(s._T_prev_full), ref.T_prev_full) < 1e-9


def test_raster_fp64_matches_oracle_1e9():
    cfg, lp = raster_multitrack(mesh=(64, 32, 16), n_steps=30)
    s, ref, its, its_ref, m, mr = _run_pair(cfg, lp, 30)
    assert its == its_ref
    assert rel(host(s.state.a), ref.a) < 1e-9


def test_single_track_fp32_within_1e4():
    cfg, lp = single_track_constant(n_steps=40, dtype="float32")
    cfg64, _ = single_track_constant(n_steps=40, dtype="float64")
    s, ctx = _solver(cfg, lp)
    ref = o.OracleSolver(make_context(cfg64, lp)).initialize()
    dt = ref.ctx.num.dt
    for n in range(40):
        s.step(n * dt, dt)
        ref.step(n * dt, dt)
    assert rel(host(s.state.a), ref.a) < 1e-4


def test_solver_contract_metrics_and_history():
    cfg = {"simulation": {"duration": 6e-6, "dt": 6e-6, "dtype": "float64"},
           "domain": {"size": [4e-4, 4e-4, 1e-4], "mesh": [32, 32, 16]},
           "material": {"rho": 7850.0, "k": 15.0, "Cp": 500.0, "L_f": 267700.0, "T_solidus": 1700.0,
                        "T_liquidus": 1800.0, "T0": 293.0, "DeltaH_LV": 7.41e6, "R_v": 150.774,
                        "Pa": 101325.0, "T_boil": 3090.0},
           "laser": {"radius": 60.0e-6, "absorptivity": 0.3, "power_nominal": 200.0}, "io": {}}
    s, ctx = _solver(cfg, FixedLaser(2e-4, 2e-4, 200.0))
    s.track_picard_history = True
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as k
    np.testing.assert_allclose(host(k.reconstruct_surface_temperature(s.state.a, s.state)), 293.0, rtol=1e-12)
    state, m = s.step(0.0, ctx.num.dt)
    assert state is s.state
    assert set(m) == {"T_surface_max", "P_laser", "n_evap_iter", "picard_converged", "picard_rel_err"}
    assert float(m["P_laser"]) == pytest.approx(0.3 * 200.0, rel=0.1)
    assert len(s.picard_history) == m["n_evap_iter"]
    off, ctx2 = _solver(cfg, FixedLaser(2e-4, 2e-4, 200.0, is_on=False))
    _, m2 = off.step(0.0, ctx2.num.dt)
    assert float(m2["P_laser"]) == 0.0
    np.testing.assert_allclose(host(k.reconstruct_volume(off.state.a, off.state)), 293.0, rtol=1e-12)
    off.set_state(np.full((16, 32, 32), 500.0))
    np.testing.assert_allclose(host(k.reconstruct_surface_temperature(off.state.a, off.state)), 500.0, rtol=1e-12)

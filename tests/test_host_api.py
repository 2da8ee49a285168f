"""Host-side layer: config parsing, units, property expressions, registry, solver selection.

Behaviour mirrors the reference's unit tests for the same layer
(tests/unit_tests/core/test_parameters.py, test_property_expr_precision.py,
test_properties.py, tests/unit_tests/backends/test_base.py,
tests/unit_tests/solvers/test_build_solver.py); no GPU is touched.
"""

import copy
import math
import os

import numpy as np
import pytest

from spectral_galerkin_heat_amd.core import units
from spectral_galerkin_heat_amd.core.laser import (RasterLaserPath, SingleTrackLaser,
                                                   build_laser_profile,
                                                   super_gaussian_area_factor)
from spectral_galerkin_heat_amd.core.parameters import SimulationContext
from spectral_galerkin_heat_amd.core.properties import (PROGRAM_BYTES, MaterialModel,
                                                        liquid_fraction)
from spectral_galerkin_heat_amd.core.property_expr import PropertyExpr

BASE = {
    "simulation": {"backend": "mi355x", "duration": 1e-4, "dt": 3e-6},
    "domain": {"size": [4e-4, 4e-4, 1e-4], "mesh": [32, 32, 16]},
    "material": {"rho": 7850.0, "k": 15.0, "Cp": 500.0, "T0": 293.0, "Pa": 101325.0,
                 "T_boil": 3090.0, "DeltaH_LV": 7.41e6, "R_v": 150.774},
    "laser": {"radius": 60e-6, "absorptivity": 0.3, "power_nominal": 200.0},
    "io": {},
}


def cfg(**over):
    c = copy.deepcopy(BASE)
    for k, v in over.items():
        c[k].update(v)
    return c


def test_step_rounding_and_dtype():
    ctx = SimulationContext.from_dict(cfg())
    assert ctx.num.n_steps == round(1e-4 / 3e-6)
    assert ctx.num.dt * ctx.num.n_steps == pytest.approx(1e-4, rel=1e-15)
    assert ctx.num.dtype is np.float32
    assert isinstance(ctx.mat.k, np.float32)
    ctx64 = SimulationContext.from_dict(cfg(simulation={"dtype": "float64"}))
    assert ctx64.num.dtype is np.float64 and isinstance(ctx64.mat.k, np.float64)
    with pytest.raises(ValueError):
        SimulationContext.from_dict(cfg(simulation={"dtype": "float16"}))


def test_unit_conversion_and_mismatch():
    c = cfg(domain={"size": {"value": [0.4, 0.4, 0.1], "unit": "mm"}},
            laser={"radius": {"value": 60, "unit": "um"}},
            material={"rho": {"value": 7.85, "unit": "g/cm**3"},
                      "h_conv_top": {"value": 10.0, "unit": "W/(m2.K)"}})
    ctx = SimulationContext.from_dict(c)
    assert ctx.geom.size.x == pytest.approx(4e-4)
    assert ctx.laser.radius == pytest.approx(60e-6)
    assert float(ctx.mat.rho) == pytest.approx(7850.0)
    assert float(ctx.mat.h_conv_top) == pytest.approx(10.0)
    with pytest.raises(units.DimensionalityError):
        SimulationContext.from_dict(cfg(laser={"radius": {"value": 60, "unit": "s"}}))
    assert units.convert(1.0, "kg/m^3", "kg / m**3") == 1.0
    assert units.convert(2.0, "mm", "m") == pytest.approx(2e-3)
    with pytest.raises(units.UnitError):
        units.parse_unit("furlong")


def test_unknown_keys_and_sections_rejected():
    with pytest.raises(ValueError, match="Unknown key"):
        SimulationContext.from_dict(cfg(material={"conductivity": 1.0}))
    bad = cfg()
    bad["solver"] = {}
    with pytest.raises(ValueError, match="Unknown config section"):
        SimulationContext.from_dict(bad)
    with pytest.raises(ValueError, match="laser.path"):
        SimulationContext.from_dict(cfg(laser={"path": {"type": "gcode", "speed": 1}}))


def test_picard_overrides_are_parsed():
    ctx = SimulationContext.from_dict(cfg(simulation={"max_picard_iter": 55, "picard_tol": 1e-6,
                                                      "picard_omega": 0.25}))
    assert (ctx.num.max_picard_iter, ctx.num.picard_tol, ctx.num.picard_omega) == (55, 1e-6, 0.25)


def test_tdep_property_requires_reference():
    mat = {"k": {"solid": "9.248 + 0.01571 * T", "liquid": "12.41 + 0.003279 * T"},
           "T_solidus": 1674.15, "T_liquidus": 1697.15}
    with pytest.raises(ValueError, match="reference"):
        SimulationContext.from_dict(cfg(material=mat))
    mat["k"]["reference"] = 24.5
    ctx = SimulationContext.from_dict(cfg(material=mat))
    assert float(ctx.mat.k) == pytest.approx(24.5)
    assert not ctx.mat.model.is_constant
    # constant branches keep the constant fast path
    same = {"k": {"solid": "15", "liquid": "15"}}
    assert SimulationContext.from_dict(cfg(material=same)).mat.model.is_constant


def test_property_expressions():
    assert PropertyExpr("42").is_constant
    e = PropertyExpr("0.5*(T + 30 - fabs(T - 30))")
    np.testing.assert_allclose(e(np.array([10.0, 50.0])), [10.0, 30.0])
    for bad in ("__import__('os')", "T.real", "T if T else 1", "max(T, 1)", "x + 1", "'a'"):
        with pytest.raises(ValueError):
            PropertyExpr(bad)
    ops, consts = PropertyExpr("8084.2 - 0.42086 * T - 3.8942e-5 * T**2").to_program()
    assert consts == [8084.2, 0.42086, 3.8942e-5, 2.0] or 2.0 not in consts


def test_liquid_fraction_and_model_program():
    assert liquid_fraction(1600.0, 1650.0, 1700.0) == 0.0
    assert liquid_fraction(1675.0, 1650.0, 1700.0) == pytest.approx(0.5)
    assert liquid_fraction(1650.0, 1650.0, 1650.0) == 1.0
    model = MaterialModel.from_config(
        {"k": {"solid": "9 + 0.01*T", "liquid": "20"}, "rho": 7900.0, "Cp": 500.0}, 1650.0, 1700.0)
    assert len(model.device_program()) == PROGRAM_BYTES
    k_ref, rho_ref, c_ref = model.recommended_references(293.0, 3090.0)
    assert rho_ref == pytest.approx(7900.0) and c_ref == pytest.approx(500.0)


def test_laser_profiles_and_paths():
    assert build_laser_profile("gaussian").order == 2.0
    assert build_laser_profile("flat_top").order == 12.0
    assert build_laser_profile("super_gaussian", 6.0).order == 6.0
    with pytest.raises(ValueError):
        build_laser_profile("donut")
    assert super_gaussian_area_factor(2.0) == pytest.approx(math.pi / 2)
    assert super_gaussian_area_factor(2000.0) == pytest.approx(math.pi, rel=0.01)
    p = build_laser_profile("gaussian")
    assert p.peak_intensity(200.0, 0.35, 50e-6 ** 2) == pytest.approx(2 * 0.35 * 200 / (math.pi * 50e-6 ** 2))
    tr = SingleTrackLaser(0.2e-3, 1.8e-3, 0.5e-3, 0.5, 200.0)
    s = tr.get_state(1e-3, 1e-5)
    assert s.is_on and s.x == pytest.approx(0.7e-3) and s.v == (0.5, 0.0)
    assert not tr.get_state(4e-3, 1e-5).is_on
    r = RasterLaserPath(3, 1e-3, 2e-3, 1e-3, 80e-6, 1.0, 250.0, 0.5e-3)
    s0, s_dwell, s1 = r.get_state(0.5e-3, 1e-5), r.get_state(1.2e-3, 1e-5), r.get_state(1.6e-3, 1e-5)
    assert s0.is_on and s0.v[0] > 0 and s0.y == pytest.approx(1e-3 - 80e-6)
    assert not s_dwell.is_on
    assert s1.is_on and s1.v[0] < 0 and s1.y == pytest.approx(1e-3)
    assert not r.get_state(10e-3, 1e-5).is_on


def test_registry_and_build_solver():
    from spectral_galerkin_heat_amd.backends import (get_backend, register_backend,
                                                     registered_backends)
    from spectral_galerkin_heat_amd.solvers import build_solver
    assert "mi355x" in registered_backends()
    with pytest.raises(ValueError, match="Unknown backend"):
        get_backend("cupy")
    with pytest.raises(ValueError, match="already registered"):
        register_backend("mi355x")(lambda: None)
    ctx = SimulationContext.from_dict(cfg(simulation={"backend": "cpu_linear"}))
    with pytest.raises(ValueError, match="Unknown backend"):
        build_solver(ctx)


def test_batch_rejects_volume_forcing():
    from spectral_galerkin_heat_amd.engine.batch import check_batchable
    from spectral_galerkin_heat_amd.workloads import make_context, single_track_constant, single_track_tdep
    c, lp = single_track_constant()
    check_batchable(make_context(c, lp))
    c, lp = single_track_tdep()
    with pytest.raises(ValueError, match="latent heat"):
        check_batchable(make_context(c, lp))


def test_melt_pool_extent_counts_cells_from_the_top():
    import torch

    from spectral_galerkin_heat_amd.physics.meltpool import melt_pool_extent
    T = torch.full((8, 10, 12), 300.0, dtype=torch.float64)
    T[5:8, 3:6, 2:9] = 2000.0
    T[6, 4, 10] = 1723.0          # on the threshold counts as melted
    assert melt_pool_extent(T, (1.0, 2.0, 3.0)) == (9 * 1.0, 3 * 2.0, 3 * 3.0)
    assert melt_pool_extent(torch.full((4, 4, 4), 300.0), (1, 1, 1)) == (0.0, 0.0, 0.0)


def test_scan_server_rejects_a_request_that_never_fits():
    """A request larger than the whole pool must raise instead of waiting forever (the
    pool is a stand-in here: admission is decided on the host)."""
    from spectral_galerkin_heat_amd.engine.serving import Request, ScanServer
    from spectral_galerkin_heat_amd.engine.batch import PoolFull

    class NoRoomPool:
        def check_admissible(self, ctx, n_steps):
            pass

        def add(self, ctx, n_steps):
            raise PoolFull("no room")

    from spectral_galerkin_heat_amd.workloads import make_context, single_track_constant
    srv = ScanServer.__new__(ScanServer)
    srv.pool, srv.precision, srv.max_steps = NoRoomPool(), "mixed", 64
    srv.steps_per_tick, srv.check_every, srv.keep_results, srv.device_steps = 8, 20, False, 0
    srv._last_refusal, srv.plain = "", None
    cfg, lp = single_track_constant(mesh=(16, 8, 4), n_steps=4)
    req = Request(rid=7, ctx=make_context(cfg, lp), n_steps=4, arrival=0.0)
    with pytest.raises(RuntimeError, match="request 7 does not fit.*no room"):
        srv.serve([req], clock=lambda: 0.0, sleep=lambda s: None)


def test_pool_admission_refuses_what_no_free_room_can_hold():
    # a cuboid deeper than the pool's max_nz (its 1-D table slot), more steps than the
    # metrics rows, or forcing outside the batched engine's scope is refused before any
    # device call, whatever the pool holds; the server refuses such a request up front
    from types import SimpleNamespace
    from spectral_galerkin_heat_amd.engine.batch import BatchPool
    from spectral_galerkin_heat_amd.engine.serving import Request, ScanServer
    from spectral_galerkin_heat_amd.workloads import (make_context, single_track_constant,
                                                      single_track_tdep)
    pool = SimpleNamespace(max_nz=16, metrics_rows=8)
    cfg, lp = single_track_constant(mesh=(32, 16, 16), n_steps=4)
    BatchPool.check_admissible(pool, make_context(cfg, lp), 8)
    deep, _ = single_track_constant(mesh=(32, 16, 24), n_steps=4)
    with pytest.raises(ValueError, match="max_nz"):
        BatchPool.check_admissible(pool, make_context(deep, lp), 4)
    with pytest.raises(ValueError, match="metrics rows"):
        BatchPool.check_admissible(pool, make_context(cfg, lp), 9)
    tdep, lp2 = single_track_tdep(mesh=(16, 8, 8), n_steps=2)
    with pytest.raises(ValueError, match="latent heat"):
        BatchPool.check_admissible(pool, make_context(tdep, lp2), 2)

    class Pool:
        def check_admissible(self, ctx, n_steps):
            BatchPool.check_admissible(pool, ctx, n_steps)

    srv = ScanServer.__new__(ScanServer)
    srv.pool, srv.max_steps, srv.plain = Pool(), 64, None
    req = Request(rid=3, ctx=make_context(deep, lp), n_steps=4, arrival=0.0)
    with pytest.raises(ValueError, match="request 3: nz = 24 exceeds"):
        srv.serve([req], clock=lambda: 0.0, sleep=lambda s: None)
    # physics outside the batched engine: refused without a plain lane
    req = Request(rid=4, ctx=make_context(tdep, lp2), n_steps=2, arrival=0.0)
    with pytest.raises(ValueError, match="latent heat"):
        srv.serve([req], clock=lambda: 0.0, sleep=lambda s: None)


@pytest.mark.parametrize("pos,n_box,expected", [(5.0, 4, (3, 7, 2)), (0.0, 4, (0, 4, 0)),
                                                (100.0, 4, (16, 20, 3)), (5.0, 30, (0, 30, 5))])
def test_box_window_reference_vectors(pos, n_box, expected):
    # tests/unit_tests/physics/test_spectral_helpers.py:32-54
    from spectral_galerkin_heat_amd.physics.spectral_helpers import _calculate_subgrid_indices
    assert _calculate_subgrid_indices(pos, 1.0, 20, n_box) == expected


def test_box_workloads_parse_with_a_fine_mesh():
    from spectral_galerkin_heat_amd.workloads import make_context, single_track_box, tdep_track_box
    ctx = make_context(*single_track_box())
    assert ctx.fine.refinement == 2 and tuple(ctx.geom.n) == (80, 48, 120)
    ctx = make_context(*tdep_track_box())
    assert ctx.fine is not None and not ctx.mat.model.is_constant


def test_runner_reads_a_reference_style_yaml_with_gcode(tmp_path):
    from spectral_galerkin_heat_amd.runner import load_context
    (tmp_path / "paths").mkdir()
    (tmp_path / "paths" / "t.gcode").write_text("G21\nG90\nG0 X0.2 Y0.15\nM3 S24\nG1 X0.4 Y0.15 F9000\nM5\n")
    (tmp_path / "c.yaml").write_text(
        "simulation: {backend: cpu, duration: {value: 2.0e-4, unit: s}, dt: {value: 4.0e-6, unit: s}}\n"
        "domain: {size: {value: [0.5e-3, 0.3e-3, 0.12e-3], unit: m}, mesh: {value: [20, 12, 16], unit: ''}}\n"
        "fine_mesh: {refinement: 2, box_size: [0.2e-3, 0.15e-3, 0.04e-3]}\n"
        "material: {name: x, rho: 7900.0, k: 20.0, Cp: 500.0, T0: 293.0}\n"
        "laser: {radius: {value: 30.0e-6, unit: m}, absorptivity: 0.3, power_nominal: {value: 24.0, unit: W},\n"
        "        path: {type: gcode, file: t.gcode}}\n"
        "io: {}\n")
    ctx = load_context(tmp_path / "c.yaml", backend="mi355x", n_steps=7)
    assert ctx.backend == "mi355x" and ctx.num.n_steps == 7 and ctx.fine.refinement == 2
    ls = ctx.laser_path.get_state(1e-4, 4e-6)
    assert ls.is_on and abs(ls.x - 0.215e-3) < 1e-6 and abs(ls.v[0] - 0.15) < 1e-3


def test_xp_namespace_matches_numpy_on_the_reference_calls():
    # the array calls the reference's solver, laser and property code make through
    # backend.xp (spectral.py, laser.py:52, properties.py:71, spectral_ops.py:158-166,
    # spectral_state.py:119), on torch tensors, against NumPy
    import torch
    from spectral_galerkin_heat_amd.backends.mi355x_backend import make_xp
    xp = make_xp("cpu")
    rng = np.random.default_rng(5)
    f = rng.standard_normal((4, 5, 6))
    t = torch.tensor(f)
    for a, b in zip(xp.gradient(t, 0.1, 0.2, 0.3), np.gradient(f, 0.1, 0.2, 0.3)):
        np.testing.assert_allclose(a.numpy(), b, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(xp.gradient(t, 0.2, axis=2).numpy(), np.gradient(f, 0.2, axis=2),
                               rtol=1e-14, atol=1e-14)
    np.testing.assert_array_equal(xp.power(-1.0, xp.arange(5, dtype=np.float64)).numpy(),
                                  np.power(-1.0, np.arange(5.0)))
    np.testing.assert_array_equal(xp.where(t >= 0.1, 1.0, 0.0).numpy(),
                                  np.where(f >= 0.1, 1.0, 0.0))
    np.testing.assert_allclose(xp.exp(-2.0 * t * t).numpy(), np.exp(-2.0 * f * f), rtol=1e-15)
    np.testing.assert_allclose(xp.cos(xp.pi * t).numpy(), np.cos(np.pi * f), rtol=1e-14,
                               atol=1e-15)
    B = rng.standard_normal((6, 6))
    np.testing.assert_allclose(xp.tensordot(t, torch.tensor(B), axes=([2], [1])).numpy(),
                               np.tensordot(f, B, axes=([2], [1])), rtol=1e-13)
    np.testing.assert_allclose(xp.einsum("zyx,Xx->zyX", t, torch.tensor(B)).numpy(),
                               np.einsum("zyx,Xx->zyX", f, B), rtol=1e-13)
    assert float(xp.vdot(t, t)) == pytest.approx(float(np.vdot(f, f)), rel=1e-14)


def test_device_out_of_memory_is_its_own_error(monkeypatch):
    # the library's SGH_ERR_NOMEM surfaces as DeviceOutOfMemory; torch's allocator error
    # and chained causes are recognised the same way
    from spectral_galerkin_heat_amd import _lib

    class Stub:
        def sgh_fake(self):
            return -4

        def sgh_fail(self):
            return -2

        def sgh_last_error(self):
            return b"hipMalloc: out of memory"

    monkeypatch.setattr(_lib, "lib", Stub())
    with pytest.raises(_lib.DeviceOutOfMemory, match="out of memory"):
        _lib.call("sgh_fake")
    with pytest.raises(_lib.SGHError) as ei:
        _lib.call("sgh_fail")
    assert not isinstance(ei.value, _lib.DeviceOutOfMemory)
    assert _lib.is_device_oom(_lib.DeviceOutOfMemory("x"))
    OOM = type("OutOfMemoryError", (RuntimeError,), {})
    try:
        try:
            raise OOM("torch")
        except OOM as inner:
            raise ValueError("wrapped") from inner
    except ValueError as outer:
        assert _lib.is_device_oom(outer)
    assert not _lib.is_device_oom(ValueError("other"))


def test_serving_pool_backs_off_after_device_out_of_memory(monkeypatch):
    from spectral_galerkin_heat_amd import _lib
    from spectral_galerkin_heat_amd.engine import serving
    tried = []

    class Pool:
        def __init__(self, precision, max_cuboids, state_capacity, surface_capacity,
                     laser_capacity, max_nz, metrics_rows):
            tried.append(state_capacity * 4 / 0.85)
            if state_capacity * 4 > (1 << 30):
                raise _lib.DeviceOutOfMemory("sgh_batch_pool_create failed (-4)")
            self.variant = None

        def set_variant(self, rows, threads, lds_bytes):
            self.variant = {"rows": rows, "threads": threads, "lds_bytes": lds_bytes}

    monkeypatch.setattr(serving, "BatchPool", Pool)
    srv = serving.ScanServer(hbm_budget_bytes=8 << 30)
    assert srv.hbm_budget_bytes == 1 << 30 and len(tried) == 4
    assert srv.variant["rows"] == -1    # mixed precision serves on the resident kernel
    # FP64 has no resident kernel; variant=None keeps the pool's per-pass default
    assert serving.ScanServer(precision="fp64", hbm_budget_bytes=1 << 30).variant is None
    assert serving.ScanServer(hbm_budget_bytes=1 << 30, variant=None).variant is None
    assert serving.ScanServer(hbm_budget_bytes=1 << 30,
                              variant=(16, 1024, 150 * 1024)).variant["rows"] == 16
    # below the smallest budget the error reaches the caller
    monkeypatch.setattr(serving.ScanServer, "MIN_BUDGET", 4 << 30)
    with pytest.raises(_lib.DeviceOutOfMemory):
        serving.ScanServer(hbm_budget_bytes=8 << 30)


def test_runner_removes_its_output_after_device_out_of_memory(monkeypatch, tmp_path):
    import yaml
    from spectral_galerkin_heat_amd import _lib, runner
    from spectral_galerkin_heat_amd.workloads import single_track_constant
    cfg, _ = single_track_constant(mesh=(16, 8, 4), n_steps=2)
    path = tmp_path / "cfg.yaml"
    path.write_text(yaml.safe_dump(cfg))
    out = tmp_path / "run"

    class Runner:
        def __init__(self, ctx, out_dir=None, checkpoint_every=0):
            self.out = out_dir

        def run(self, resume=None):
            os.makedirs(self.out, exist_ok=True)
            raise _lib.DeviceOutOfMemory("sgh_batch_create failed (-4): out of memory")

    monkeypatch.setattr(runner, "HeatRunner", Runner)
    with pytest.raises(SystemExit) as ei:
        runner.main([str(path), "--out", str(out)])
    assert ei.value.code == 2 and not out.exists()

"""Continuous batching: pool admission/removal and the request loop on the MI355X.

Each served result must equal the same request run alone (bit for bit: a cuboid's
arithmetic does not depend on the batch it shares), every re-solved request must agree
with FP64 to 1e-4, and admission must respect the pool's memory capacity.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spectral_galerkin_heat_amd.workloads import batch_uneven_cuboids, make_context  # noqa: E402


def _requests(n, seed, steps_lo=10, steps_hi=40, gap=0.004, burst=1):
    """n requests; the first `burst` arrive together at t = 0, the rest `gap` s apart."""
    from spectral_galerkin_heat_amd.engine.serving import Request
    rng = np.random.default_rng(seed)
    cases = batch_uneven_cuboids(n=n, seed=seed, n_steps=steps_hi)
    out = []
    for i, (cfg, lp) in enumerate(cases):
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        t = max(0, i - burst + 1) * gap
        out.append(Request(rid=i, ctx=make_context(cfg, lp), n_steps=steps, arrival=t))
    return out


def test_pool_add_remove_and_capacity():
    from spectral_galerkin_heat_amd.engine.batch import BatchPool, PoolFull
    (cfg, lp), = batch_uneven_cuboids(n=1, seed=1, n_steps=5)
    ctx = make_context(cfg, lp)
    N = int(np.prod(tuple(ctx.geom.n)))
    pool = BatchPool(precision="mixed", max_cuboids=4, state_capacity=2 * N + 64,
                     surface_capacity=1 << 16, laser_capacity=64, max_nz=32, metrics_rows=8)
    a = pool.add(ctx, 5)
    b = pool.add(ctx, 5)
    with pytest.raises(PoolFull):
        pool.add(ctx, 5)
    pool.run(5)
    pool.synchronize()
    assert torch.equal(pool.state(a), pool.state(b))
    pool.remove(a)
    c = pool.add(ctx, 5)
    pool.run(5)
    pool.synchronize()
    assert torch.equal(pool.state(c), pool.state(b))
    pool.close()


def test_pool_refuses_a_cuboid_deeper_than_max_nz():
    # the 1-D table slot of a cuboid holds 2·(512 + max_nz) doubles: a deeper cuboid would
    # write into the next slot's tables, so both the binding and the C entry point refuse it
    import ctypes as C
    from spectral_galerkin_heat_amd import _lib
    from spectral_galerkin_heat_amd.engine.batch import BatchPool, cuboid_spec
    from spectral_galerkin_heat_amd.workloads import single_track_constant
    pool = BatchPool(precision="mixed", max_cuboids=2, state_capacity=1 << 20,
                     surface_capacity=1 << 14, laser_capacity=64, max_nz=16, metrics_rows=8)
    ok, lp = single_track_constant(mesh=(32, 16, 16), n_steps=4, dtype="float32")
    deep, _ = single_track_constant(mesh=(32, 16, 24), n_steps=4, dtype="float32")
    with pytest.raises(ValueError, match="max_nz"):
        pool.add(make_context(deep, lp), 4)
    sp, tab = cuboid_spec(make_context(deep, lp), 4, _lib.F32)
    slot = C.c_int(-1)
    rc = _lib.lib.sgh_batch_add(pool._plan, C.byref(sp), C.byref(slot), pool.stream.cuda_stream)
    assert rc == -1 and "max_nz" in _lib.last_error()
    i = pool.add(make_context(ok, lp), 4)
    pool.run(4)
    pool.synchronize()
    assert pool.steps()[i] == 4
    pool.close()


def test_pool_releases_matrices_of_departed_mode_counts():
    # 40 distinct surface widths come and go one after another; each width's cosine
    # matrices are freed with its last cuboid, so the table never fills
    from spectral_galerkin_heat_amd.engine.batch import BatchPool
    from spectral_galerkin_heat_amd.workloads import single_track_constant
    pool = BatchPool(precision="fp64", max_cuboids=4, state_capacity=1 << 20,
                     surface_capacity=1 << 14, laser_capacity=256, max_nz=8, metrics_rows=4)
    keep_cfg, lp = single_track_constant(mesh=(24, 20, 4), n_steps=2)
    keep = pool.add(make_context(keep_cfg, lp), 2)
    for nx in range(200, 240):
        cfg, lp2 = single_track_constant(mesh=(nx, 12, 4), n_steps=2)
        i = pool.add(make_context(cfg, lp2), 2)
        assert pool.matrix_sizes() == sorted({12, 20, 24, nx})
        pool.run(1)
        pool.remove(i)
        assert pool.matrix_sizes() == [20, 24]
    pool.remove(keep)
    assert pool.matrix_sizes() == []
    pool.close()


def test_served_results_equal_solo_runs():
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    from spectral_galerkin_heat_amd.engine.serving import ScanServer
    reqs = _requests(7, seed=21, burst=5)
    srv = ScanServer(precision="mixed", max_cuboids=4, hbm_budget_bytes=1 << 28, max_steps=64,
                     steps_per_tick=3)
    rep = srv.serve(reqs)
    assert len(rep.requests) == 7
    # five arrive together for four slots: one joins only after the pool has stepped
    assert any(r.start_step > 0 for r in rep.requests)
    for r in rep.requests:
        solo = BatchEngine([r.ctx], precision="mixed", n_steps=r.n_steps, variant=srv.variant)
        solo.run(r.n_steps)
        solo.synchronize()
        assert torch.equal(solo.state(0), r.result), r.rid
        np.testing.assert_array_equal(solo.metrics()[:, 0], r.metrics)
        solo.close()
    s = rep.summary()
    assert s["p99_latency_s"] >= s["p50_latency_s"] > 0
    srv.close()


def test_fp64_recheck_within_1e4():
    from spectral_galerkin_heat_amd.engine.serving import ScanServer
    reqs = _requests(3, seed=5, steps_lo=20, steps_hi=30)
    srv = ScanServer(precision="mixed", max_cuboids=8, hbm_budget_bytes=1 << 28, max_steps=64,
                     check_every=1)
    rep = srv.serve(reqs)
    checks = srv.fp64_checks(rep)
    assert len(checks) == 3 and max(checks) < 1e-4, checks
    srv.close()


def test_c5_poisson_stream_with_rasters():
    """C5: 60 requests from the serving distribution (3 of them 256×128×64 rasters with
    1-5 tracks, 200-2000 steps each) arriving as a 40 Hz Poisson stream. Every served
    result equals a solo run of the same request bit for bit for a sample holding a
    raster, and every 20th request re-solved in FP64 agrees to 1e-4."""
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    from spectral_galerkin_heat_amd.engine.serving import Request, ScanServer
    from spectral_galerkin_heat_amd.workloads import serving_stream
    stream = serving_stream(60, 40.0, seed=3)
    reqs = [Request(rid=i, ctx=make_context(cfg, lp), n_steps=n, arrival=t)
            for i, (cfg, lp, n, t) in enumerate(stream)]
    rasters = [r.rid for r in reqs if tuple(r.ctx.geom.n) == (256, 128, 64)]
    assert len(rasters) >= 2
    srv = ScanServer(precision="mixed", max_cuboids=160, hbm_budget_bytes=8 << 30,
                     max_steps=2048, steps_per_tick=8, check_every=20)
    rep = srv.serve(reqs)
    assert len(rep.requests) == 60
    assert rep.peak_active >= 10
    s = rep.summary()
    print("C5 serving:", {k: s[k] for k in ("p50_latency_s", "p99_latency_s",
                                            "aggregate_mode_steps_per_s", "peak_active")})
    by = {r.rid: r for r in rep.requests}
    for rid in (rasters[0], 7, 33):
        r = by[rid]
        solo = BatchEngine([r.ctx], precision="mixed", n_steps=r.n_steps, variant=srv.variant)
        solo.run(r.n_steps)
        solo.synchronize()
        assert torch.equal(solo.state(0), r.result), rid
        np.testing.assert_array_equal(solo.metrics()[:, 0], r.metrics)
        solo.close()
    checks = srv.fp64_checks(rep)
    assert len(checks) == 3 and max(checks) < 1e-4, checks
    srv.close()


def test_serving_mixes_plain_lane_requests_and_each_equals_its_solo_run():
    # latent-heat, convection and T-dependent requests go to the plain lane beside the
    # batched pool; served results equal the solo runs bit for bit, and the FP64 re-check
    # of a plain request holds 1e-4
    from spectral_galerkin_heat_amd.backends import get_backend
    from spectral_galerkin_heat_amd.engine.serving import Request, ScanServer
    from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
    from spectral_galerkin_heat_amd.workloads import single_track_box, single_track_tdep
    reqs = _requests(6, seed=21, steps_lo=6, steps_hi=14, gap=0.003, burst=2)
    plain_cases = [single_track_tdep(mesh=(32, 16, 8), n_steps=8, dtype="float32"),
                   single_track_box(mesh=(40, 24, 60), n_steps=5,
                                    box_size=(0.3e-3, 0.3e-3, 0.05e-3)),
                   single_track_tdep(mesh=(48, 32, 16), n_steps=7)]
    for j, (cfg, lp) in enumerate(plain_cases):
        sim = cfg["simulation"]
        steps = int(round(sim["duration"] / sim["dt"]))
        reqs.append(Request(rid=100 + 20 * j, ctx=make_context(cfg, lp), n_steps=steps,
                            arrival=0.002 * j))
    srv = ScanServer(precision="mixed", max_cuboids=8, hbm_budget_bytes=256 << 20,
                     max_steps=64, check_every=20, plain_slots=2)
    rep = srv.serve(reqs)
    srv.fp64_checks(rep)
    summ = rep.summary()
    assert summ["requests"] == len(reqs) and summ["plain_requests"] == 3
    for r in rep.requests:
        if not r.plain:
            continue
        s = SpectralSolver(get_backend("mi355x"), r.ctx)
        s.initialize()
        dt = r.ctx.num.dt
        for k in range(r.n_steps):
            s.step(k * dt, dt)
        torch.cuda.synchronize()
        assert torch.equal(s.state.a, r.result), r.rid
        assert r.metrics.shape == (r.n_steps, 5)
        print(f"plain request {r.rid} {tuple(r.result.shape)}: identical to its solo run, "
              f"latency {r.latency:.3f} s, FP64 re-check {r.fp64_rel_l2:.2e}")
        if not np.isnan(r.fp64_rel_l2):
            assert r.fp64_rel_l2 < 1e-4
    srv.close()

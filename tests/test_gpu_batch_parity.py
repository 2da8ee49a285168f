"""GPU parity of the batched surface-collapsed engine against the FP64 oracle.

FP64 engine: relative L2 of each cuboid's modal state ≤ 1e-9 against the oracle with
the same per-step Picard counts. Mixed engine (FP32 state and transforms, FP64
tables/norms): ≤ 1e-4. A cuboid gives bit-identical results alone and inside a batch,
and a graph-replayed step equals the eagerly launched one.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.sgh_oracle import OracleSolver, modal_rel_l2, modal_rel_l2_excursion  # noqa: E402
from spectral_galerkin_heat_amd.workloads import (batch_uneven_cuboids, make_context,  # noqa: E402
                                                  raster_multitrack, single_track_constant)


def _engine(cases, precision, **kw):
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    return BatchEngine([make_context(c, lp) for c, lp in cases], precision=precision, **kw)


def _oracle_run(cfg, lp, n):
    ref = OracleSolver(make_context(cfg, lp)).initialize()
    dt = ref.ctx.num.dt
    its, tmax, pl = [], [], []
    for k in range(n):
        _, m = ref.step(k * dt, dt)
        its.append(m["n_evap_iter"])
        tmax.append(m["T_surface_max"])
        pl.append(m["P_laser"])
    return ref, its, tmax, pl


def test_batch_fp64_matches_oracle_per_cuboid():
    cases = batch_uneven_cuboids(n=4, seed=5, n_steps=15)
    eng = _engine(cases, "fp64")
    eng.run(15)
    eng.synchronize()
    met = eng.metrics()
    for i, (cfg, lp) in enumerate(cases):
        ref, its, tmax, pl = _oracle_run(cfg, lp, 15)
        a = eng.state(i).cpu().numpy()
        assert modal_rel_l2(a, ref.a) < 1e-9, i
        assert met[:, i, 2].astype(int).tolist() == its
        np.testing.assert_allclose(met[:, i, 1], tmax, rtol=1e-9)
        np.testing.assert_allclose(met[:, i, 0], pl, rtol=1e-12)
    eng.close()


def test_batch_single_track_320_steps_fp64():
    cfg, lp = single_track_constant()
    eng = _engine([(cfg, lp)], "fp64")
    eng.run(320)
    eng.synchronize()
    ref, its, tmax, _ = _oracle_run(cfg, lp, 320)
    assert modal_rel_l2(eng.state(0).cpu().numpy(), ref.a) < 1e-9
    assert eng.metrics()[:, 0, 2].astype(int).tolist() == its
    eng.close()


def test_batch_raster_with_dwell_fp64():
    cfg, lp = raster_multitrack(mesh=(64, 32, 16), n_steps=60, size=(1.0e-3, 0.5e-3, 0.25e-3))
    eng = _engine([(cfg, lp)], "fp64")
    eng.run(60)
    eng.synchronize()
    ref, its, _, _ = _oracle_run(cfg, lp, 60)
    assert modal_rel_l2(eng.state(0).cpu().numpy(), ref.a) < 1e-9
    eng.close()


def test_batch_mixed_within_1e4():
    cases = batch_uneven_cuboids(n=6, seed=2, n_steps=40)
    eng = _engine(cases, "mixed")
    eng.run(40)
    eng.synchronize()
    for i, (cfg, lp) in enumerate(cases):
        ref, _, _, _ = _oracle_run(cfg, lp, 40)
        a = eng.state(i).double().cpu().numpy()
        assert modal_rel_l2(a, ref.a) < 1e-4, i
        g = ref.ctx.geom
        mean0 = float(ref.ctx.mat.T0) * np.sqrt(g.size.x * g.size.y * g.size.z)
        print(f"cuboid {i}: rel-L2 {modal_rel_l2(a, ref.a):.2e}, "
              f"(T-T0) rel-L2 {modal_rel_l2_excursion(a, ref.a, mean0):.2e}")
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_alone_equals_in_mix(precision):
    cases = batch_uneven_cuboids(n=5, seed=9, n_steps=12)
    mix = _engine(cases, precision)
    mix.run(12)
    mix.synchronize()
    for i in (0, 3):
        alone = _engine([cases[i]], precision)
        alone.run(12)
        alone.synchronize()
        assert torch.equal(alone.state(0), mix.state(i))
        np.testing.assert_array_equal(alone.metrics()[:, 0], mix.metrics()[:, i])
        alone.close()
    mix.close()


def test_graph_replay_equals_eager_launches():
    cases = batch_uneven_cuboids(n=3, seed=4, n_steps=9)
    g = _engine(cases, "mixed")
    e = _engine(cases, "mixed")
    for _ in range(3):
        g.run(3, use_graph=True)
        e.run(3, use_graph=False)
        g.synchronize()
        e.synchronize()
        for i in range(3):
            assert torch.equal(g.state(i), e.state(i))
    g.close()
    e.close()


@pytest.mark.parametrize("lds", [150 * 1024, 159 * 1024, 72 * 1024])
def test_every_surface_variant_matches_oracle(lds):
    """Each (rows, threads) variant and LDS budget keeps the engines on the oracle. Small
    budgets stage the operands in chunks; a budget that cannot hold a 16-column chunk of a
    256-row operand next to the item's rows is refused (some FP64 and 72 KiB cases)."""
    from spectral_galerkin_heat_amd import _lib
    from spectral_galerkin_heat_amd.engine.autotune import surface_variants
    cases = batch_uneven_cuboids(n=3, seed=21, n_steps=5)
    refs = [_oracle_run(cfg, lp, 5) for cfg, lp in cases]
    for precision in ("fp64", "mixed"):
        eng = _engine(cases, precision)
        for rows, threads in surface_variants():
            try:
                eng.set_variant(rows, threads, lds)
            except _lib.SGHError:
                assert precision == "fp64" or lds < 100 * 1024
                continue
            eng.rewind()
            eng.run(5)
            eng.synchronize()
            met = eng.metrics()
            for i, (ref, its, _, _) in enumerate(refs):
                a = eng.state(i).double().cpu().numpy()
                tol = 1e-9 if precision == "fp64" else 1e-4
                assert modal_rel_l2(a, ref.a) < tol, (precision, rows, threads, lds, i)
                if precision == "fp64":
                    assert met[:, i, 2].astype(int).tolist() == its
        eng.close()


def test_autotuned_engine_runs_the_chosen_variant():
    from spectral_galerkin_heat_amd.engine.autotune import autotune
    cases = batch_uneven_cuboids(n=8, seed=3, n_steps=6)
    ctxs = [make_context(c, lp) for c, lp in cases]
    res = autotune(ctxs, "mixed", steps=2, reps=1)
    assert len(res.table) >= 4
    assert res.ms_per_step == min(r["ms_per_step"] for r in res.table)
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    eng = BatchEngine(ctxs, precision="mixed", variant=res.variant)
    v = eng.variant
    assert (v["rows"], v["threads"], v["lds_bytes"]) == res.variant
    eng.run(6)
    eng.synchronize()
    ref = BatchEngine(ctxs, precision="fp64")
    ref.run(6)
    ref.synchronize()
    for i in range(len(ctxs)):
        assert modal_rel_l2(eng.state(i).double().cpu().numpy(), ref.state(i).cpu().numpy()) < 1e-4
    eng.close()
    ref.close()


def test_alone_equals_in_mix_for_a_tuned_variant():
    cases = batch_uneven_cuboids(n=4, seed=17, n_steps=6)
    mix = _engine(cases, "mixed", variant=(32, 512, 150 * 1024))
    mix.run(6)
    mix.synchronize()
    alone = _engine(cases[2:3], "mixed", variant=(32, 512, 150 * 1024))
    alone.run(6)
    alone.synchronize()
    assert torch.equal(mix.state(2), alone.state(0))
    mix.close()
    alone.close()


def test_c4_sixty_four_cuboids_fp64_matches_oracle():
    """The benchmark batch itself: 64 uneven cuboids from default_rng(0), 3 steps of the
    FP64 engine against the oracle per cuboid (≤ 1e-9, identical Picard counts). Its work
    list holds more surface items than the 256 CUs, so a second round of workgroups runs,
    and its cuboids are binned over the 8 XCDs."""
    cases = batch_uneven_cuboids(n=64, seed=0, n_steps=3)
    eng = _engine(cases, "fp64")
    eng.run(3)
    eng.synchronize()
    info = eng.info            # the work lists are built by the first run
    assert info["surface_items"] > 256, info
    met = eng.metrics()
    worst = 0.0
    for i, (cfg, lp) in enumerate(cases):
        ref, its, tmax, pl = _oracle_run(cfg, lp, 3)
        err = modal_rel_l2(eng.state(i).cpu().numpy(), ref.a)
        worst = max(worst, err)
        assert err < 1e-9, (i, err)
        assert met[:, i, 2].astype(int).tolist() == its, i
        np.testing.assert_allclose(met[:, i, 1], tmax, rtol=1e-9)
    print(f"C4 64 cuboids, 3 FP64 steps: worst rel-L2 vs oracle {worst:.2e}")
    eng.close()


def test_c4_mixed_autotuned_matches_fp64_over_1000_steps():
    """The full C4 run: 1000 steps of the mixed engine with the autotuned launch shape
    against the FP64 engine (≤ 1e-4 relative L2 per cuboid; melt-pool width and depth of
    the cell-centred fields within one grid spacing)."""
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    from spectral_galerkin_heat_amd.physics.meltpool import cell_temperature, melt_pool_extent
    cases = batch_uneven_cuboids(n=64, seed=0, n_steps=1000)
    ctxs = [make_context(c, lp) for c, lp in cases]
    mixed = BatchEngine(ctxs, precision="mixed", variant="auto")
    fp64 = BatchEngine(ctxs, precision="fp64")
    mixed.run(1000)
    fp64.run(1000)
    mixed.synchronize()
    fp64.synchronize()
    worst, melted = 0.0, 0
    for i, ctx in enumerate(ctxs):
        a = mixed.state(i).double()
        b = fp64.state(i)
        err = modal_rel_l2(a.cpu().numpy(), b.cpu().numpy())
        worst = max(worst, err)
        assert err < 1e-4, (i, err)
        g = ctx.geom
        d = (g.size.x / g.n.x, g.size.y / g.n.y, g.size.z / g.n.z)
        ea = melt_pool_extent(cell_temperature(a, g), d)
        eb = melt_pool_extent(cell_temperature(b, g), d)
        melted += eb[0] > 0
        assert abs(ea[1] - eb[1]) <= d[1] * (1 + 1e-9), (i, ea, eb)
        assert abs(ea[2] - eb[2]) <= d[2] * (1 + 1e-9), (i, ea, eb)
    print(f"C4 1000 steps, variant {mixed.variant}: worst mixed-vs-FP64 rel-L2 {worst:.2e}, "
          f"{melted} cuboids melt at the end")
    mixed.close()
    fp64.close()

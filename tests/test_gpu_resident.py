"""GPU checks of the resident step kernel (batched engine variant rows = -1).

One workgroup per cuboid runs the whole surface part of a step with split-bf16 MFMA
transforms and FP32 fields. Checked: each cuboid against the FP64 oracle (≤ 1e-4
relative L2), alone equals in a mix (bit for bit), graph replay equals eager launches,
a surface wider than 128 runs on per-pass launches beside the resident ones, and the
64-cuboid benchmark batch over 1000 steps against the FP64 engine (≤ 1e-4, melt-pool
width and depth within one cell).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.sgh_oracle import OracleSolver, modal_rel_l2, modal_rel_l2_excursion  # noqa: E402
from spectral_galerkin_heat_amd.workloads import (batch_uneven_cuboids, make_context,  # noqa: E402
                                                  raster_multitrack, single_track_constant)

RESIDENT = (-1, 512, 150 * 1024)


def _engine(cases, precision="mixed", variant=RESIDENT, **kw):
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    return BatchEngine([make_context(c, lp) for c, lp in cases], precision=precision,
                       variant=variant, **kw)


def _oracle(cfg, lp, n):
    ref = OracleSolver(make_context(cfg, lp)).initialize()
    dt = ref.ctx.num.dt
    its = []
    for k in range(n):
        _, m = ref.step(k * dt, dt)
        its.append(m["n_evap_iter"])
    return ref, its


def test_resident_variant_is_selected_and_reported():
    eng = _engine(batch_uneven_cuboids(n=2, seed=1, n_steps=2))
    v = eng.variant
    assert (v["rows"], v["threads"]) == (-1, 512)
    eng.set_timing(True)
    eng.run(2)
    eng.synchronize()
    t = eng.timing()
    assert t["surface_launches"] == 2 and t["volume_launches"] == 2, t   # one launch per step
    eng.close()
    from spectral_galerkin_heat_amd import _lib
    with pytest.raises(_lib.SGHError, match="mixed precision"):
        _engine(batch_uneven_cuboids(n=1, seed=1, n_steps=2), precision="fp64")


def test_resident_matches_oracle_per_cuboid():
    cases = batch_uneven_cuboids(n=6, seed=2, n_steps=40)
    eng = _engine(cases)
    eng.run(40)
    eng.synchronize()
    met = eng.metrics()
    for i, (cfg, lp) in enumerate(cases):
        ref, its = _oracle(cfg, lp, 40)
        a = eng.state(i).double().cpu().numpy()
        g = ref.ctx.geom
        mean0 = float(ref.ctx.mat.T0) * np.sqrt(g.size.x * g.size.y * g.size.z)
        err, exc = modal_rel_l2(a, ref.a), modal_rel_l2_excursion(a, ref.a, mean0)
        print(f"cuboid {i} {tuple(g.n)}: rel-L2 {err:.2e}, (T-T0) {exc:.2e}, "
              f"iterations {met[:, i, 2].astype(int).tolist()[-5:]} vs {its[-5:]}")
        assert err < 1e-4, (i, err)
        # FP32 resolves a relative residual of about 6e-8, so a 1e-8 tolerance is met or
        # missed by rounding in single steps; the counts agree on average
        assert abs(met[:, i, 2].mean() - np.mean(its)) <= 1.0
        assert np.all(met[:, i, 0] > 0)
    eng.close()


def test_resident_single_track_320_steps():
    cfg, lp = single_track_constant()
    eng = _engine([(cfg, lp)])
    eng.run(320)
    eng.synchronize()
    ref, its = _oracle(cfg, lp, 320)
    a = eng.state(0).double().cpu().numpy()
    err = modal_rel_l2(a, ref.a)
    print(f"C1 320 steps resident: rel-L2 vs oracle {err:.2e}")
    assert err < 1e-4
    eng.close()


def test_resident_alone_equals_in_mix_and_graph_equals_eager():
    cases = batch_uneven_cuboids(n=5, seed=9, n_steps=12)
    mix = _engine(cases)
    mix.run(12)
    mix.synchronize()
    for i in (0, 3):
        alone = _engine([cases[i]])
        alone.run(12, use_graph=False)
        alone.synchronize()
        assert torch.equal(alone.state(0), mix.state(i))
        np.testing.assert_array_equal(alone.metrics()[:, 0], mix.metrics()[:, i])
        alone.close()
    mix.close()


def test_wide_surface_runs_per_pass_beside_the_resident_cuboids():
    # a surface wider than 128 runs on per-pass launches in the same step as the resident
    # launch of the others: each cuboid matches the oracle, and each equals its run alone
    # (the narrow ones on the resident kernel, the wide one on per-pass launches)
    cfg, lp = raster_multitrack(mesh=(160, 32, 16), n_steps=20, size=(2.0e-3, 0.5e-3, 0.25e-3))
    cases = batch_uneven_cuboids(n=2, seed=3, n_steps=20) + [(cfg, lp)]
    eng = _engine(cases)
    eng.set_timing(True)
    eng.run(2)
    eng.synchronize()
    t = eng.timing()
    eng.set_timing(False)
    teams = eng.team_sizes
    assert teams[0] >= 1 and teams[1] >= 1 and teams[2] == 0, teams
    assert t["surface_launches"] > 2 and t["volume_launches"] == 2, t
    eng.run(18)
    eng.synchronize()
    # the resident launch on its side stream, eager instead of graph-replayed
    eager = _engine(cases)
    eager.run(20, use_graph=False)
    eager.synchronize()
    for i in range(len(cases)):
        assert torch.equal(eager.state(i), eng.state(i)), i
    eager.close()
    for i, (c, l) in enumerate(cases):
        ref, _ = _oracle(c, l, 20)
        assert modal_rel_l2(eng.state(i).double().cpu().numpy(), ref.a) < 1e-4, i
        alone = _engine([cases[i]])
        alone.run(20)
        alone.synchronize()
        assert torch.equal(alone.state(0), eng.state(i)), i
        np.testing.assert_array_equal(alone.metrics()[:, 0], eng.metrics()[:, i])
        alone.close()
    eng.close()


def test_resident_c4_1000_steps_against_fp64_engine():
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    from spectral_galerkin_heat_amd.physics.meltpool import cell_temperature, melt_pool_extent
    cases = batch_uneven_cuboids(n=64, seed=0, n_steps=1000)
    ctxs = [make_context(c, lp) for c, lp in cases]
    res = BatchEngine(ctxs, precision="mixed", variant=RESIDENT)
    ref = BatchEngine(ctxs, precision="fp64")
    res.run(1000)
    ref.run(1000)
    res.synchronize()
    ref.synchronize()
    worst = worst_exc = 0.0
    for i, ctx in enumerate(ctxs):
        a = res.state(i).double()
        b = ref.state(i)
        g = ctx.geom
        mean0 = float(ctx.mat.T0) * np.sqrt(g.size.x * g.size.y * g.size.z)
        err = modal_rel_l2(a.cpu().numpy(), b.cpu().numpy())
        worst = max(worst, err)
        worst_exc = max(worst_exc, modal_rel_l2_excursion(a.cpu().numpy(), b.cpu().numpy(), mean0))
        assert err < 1e-4, (i, err)
        d = (g.size.x / g.n.x, g.size.y / g.n.y, g.size.z / g.n.z)
        ea = melt_pool_extent(cell_temperature(a, g), d)
        eb = melt_pool_extent(cell_temperature(b, g), d)
        assert abs(ea[1] - eb[1]) <= d[1] * (1 + 1e-9), (i, ea, eb)
        assert abs(ea[2] - eb[2]) <= d[2] * (1 + 1e-9), (i, ea, eb)
    print(f"C4 1000 steps resident vs FP64: worst rel-L2 {worst:.2e}, worst (T-T0) {worst_exc:.2e}")
    res.close()
    ref.close()


def test_team_sizes_give_identical_results():
    # the 64-cuboid batch: the largest surfaces run on teams of several workgroups; with the
    # teams capped at 1 or 3 the states and metrics must be the same bit for bit
    cases = batch_uneven_cuboids(n=64, seed=0, n_steps=30)
    runs = {}
    for cap in (8, 1, 3):
        eng = _engine(cases)
        eng.set_team_max(cap)
        eng.run(30)
        eng.synchronize()
        sizes = eng.team_sizes
        runs[cap] = ([eng.state(i).clone() for i in range(eng.n)], eng.metrics(), sizes)
        print(f"team_max {cap}: team sizes {sorted(sizes, reverse=True)[:12]}..., "
              f"{sum(sizes)} workgroups")
        eng.close()
    assert max(runs[8][2]) > 1 and max(runs[1][2]) == 1 and max(runs[3][2]) <= 3
    assert sum(runs[8][2]) <= 256
    for cap in (1, 3):
        for i, (a, b) in enumerate(zip(runs[8][0], runs[cap][0])):
            assert torch.equal(a, b), (cap, i, float((a - b).abs().max()))
        np.testing.assert_array_equal(runs[8][1], runs[cap][1])


def test_large_cuboid_alone_on_a_team_matches_oracle():
    cfg, lp = raster_multitrack(mesh=(128, 128, 16), n_steps=30, size=(2.0e-3, 2.0e-3, 0.25e-3))
    eng = _engine([(cfg, lp)])
    eng.run(30)
    eng.synchronize()
    assert eng.team_sizes[0] == 8, eng.team_sizes
    ref, its = _oracle(cfg, lp, 30)
    err = modal_rel_l2(eng.state(0).double().cpu().numpy(), ref.a)
    met = eng.metrics()
    print(f"128x128x16 on a team of 8, 30 steps: rel-L2 vs oracle {err:.2e}, iterations "
          f"{met[:, 0, 2].astype(int).tolist()[-5:]} vs {its[-5:]}")
    assert err < 1e-4
    assert abs(met[:, 0, 2].mean() - np.mean(its)) <= 1.0
    single = _engine([(cfg, lp)])
    single.set_team_max(1)
    single.run(30)
    single.synchronize()
    assert single.team_sizes[0] == 1
    assert torch.equal(single.state(0), eng.state(0))
    np.testing.assert_array_equal(single.metrics(), met)
    single.close()
    eng.close()

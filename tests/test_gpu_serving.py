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
        solo = BatchEngine([r.ctx], precision="mixed", n_steps=r.n_steps)
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

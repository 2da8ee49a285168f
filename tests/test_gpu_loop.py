"""The persistent step kernel of the batched engine against the per-pass launches.

With ``set_loop(True)`` each step's surface work (warm start, every Picard iteration,
step end) runs as one launch whose work items meet at per-cuboid device-scope barriers.
It must give the iterates of the per-pass launches of the same variant: identical Picard
counts and metrics, identical modal states; and, in FP64, the oracle's ≤ 1e-9. Work
lists with more items than CUs, and variants without a persistent kernel, fall back to the
per-pass launches.
"""

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="the persistent step kernel is switched off: a GPU memory "
                                      "fault in a test run that used it is not yet explained")]

torch = pytest.importorskip("torch")

from oracle.sgh_oracle import OracleSolver, modal_rel_l2  # noqa: E402
from spectral_galerkin_heat_amd.core.laser import SingleTrackLaser  # noqa: E402
from spectral_galerkin_heat_amd.workloads import (MATERIAL_316L, _cfg,  # noqa: E402
                                                  batch_uneven_cuboids, make_context)


def _engine(cases, precision, variant):
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    return BatchEngine([make_context(c, lp) for c, lp in cases], precision=precision,
                       variant=variant)


def _run(eng, n, loop, use_graph=True):
    eng.rewind()
    eng.set_loop(loop)
    eng.run(n, use_graph=use_graph)
    eng.synchronize()
    states = [eng.state(i).clone() for i in range(eng.n)]
    return states, eng.metrics(), eng.loop_state


def _same(sa, ma, sb, mb):
    for i, (a, b) in enumerate(zip(sa, sb)):
        assert torch.equal(a, b), (i, float((a - b).abs().max()))
    np.testing.assert_array_equal(ma[:, :, 2], mb[:, :, 2])     # Picard counts
    np.testing.assert_array_equal(ma, mb)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("threads", [512, 1024])
def test_persistent_step_equals_per_pass_launches(precision, threads):
    cases = batch_uneven_cuboids(n=20, seed=3, n_steps=5)
    eng = _engine(cases, precision, (16, threads, 153600))
    sa, ma, la = _run(eng, 5, False)
    sb, mb, lb = _run(eng, 5, True)
    assert not la["active"] and lb["active"], (la, lb, eng.info)
    _same(sa, ma, sb, mb)
    # eagerly launched steps of the persistent kernel equal the graph-replayed ones
    sc, mc, lc = _run(eng, 5, True, use_graph=False)
    assert lc["active"]
    _same(sb, mb, sc, mc)
    eng.close()


def _split_case(mesh, n_steps):
    size = (1.0e-3 * mesh[0] / 32 + 0.2e-3, 0.6e-3 * mesh[1] / 32 + 0.2e-3, 0.3e-3)
    cfg = _cfg(size, mesh, 5e-6, n_steps, MATERIAL_316L, 40e-6, 0.35, 300.0, 1e-8, 20,
               "float64")
    return cfg, SingleTrackLaser(0.2 * size[0], 0.8 * size[0], 0.5 * size[1], 1.0, 300.0)


def test_persistent_step_with_column_split_items_fp64_matches_oracle():
    # surfaces wider than 128 are split into column blocks; ragged and tiny surfaces too
    meshes = [(144, 20, 4), (256, 48, 4), (33, 17, 5), (2, 3, 2), (130, 70, 3)]
    n = 3
    cases = [_split_case(m, n) for m in meshes]
    eng = _engine(cases, "fp64", (16, 512, 153600))
    sa, ma, _ = _run(eng, n, False)
    sb, mb, lb = _run(eng, n, True)
    assert lb["active"]
    _same(sa, ma, sb, mb)
    for i, (cfg, lp) in enumerate(cases):
        ref = OracleSolver(make_context(cfg, lp)).initialize()
        dt = ref.ctx.num.dt
        its = [ref.step(k * dt, dt)[1]["n_evap_iter"] for k in range(n)]
        assert modal_rel_l2(sb[i].cpu().numpy(), ref.a) < 1e-9, meshes[i]
        assert mb[:, i, 2].astype(int).tolist() == its, meshes[i]
    eng.close()


def test_cuboids_finishing_at_different_steps():
    # cuboids with 2, 4 and 6 steps: finished ones leave the persistent kernel at once
    cases = batch_uneven_cuboids(n=6, seed=11, n_steps=6)
    for i, (cfg, _) in enumerate(cases):
        cfg["simulation"]["duration"] = cfg["simulation"]["dt"] * (2 + 2 * (i % 3))
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    eng = BatchEngine([make_context(c, lp) for c, lp in cases], precision="mixed",
                      variant=(16, 512, 153600))
    sa, ma, _ = _run(eng, 6, False)
    sb, mb, lb = _run(eng, 6, True)
    assert lb["active"]
    _same(sa, ma, sb, mb)
    eng.close()


def test_work_list_beyond_the_cus_falls_back_to_per_pass_launches():
    # 64 uneven cuboids in 16-row items: about 290 items > 256 CUs
    cases = batch_uneven_cuboids(n=64, seed=0, n_steps=2)
    eng = _engine(cases, "mixed", (16, 512, 153600))
    sa, ma, _ = _run(eng, 2, False)
    sb, mb, lb = _run(eng, 2, True)
    assert lb["requested"] and not lb["active"], (lb, eng.info)
    _same(sa, ma, sb, mb)
    eng.close()


def test_variants_without_a_persistent_kernel_fall_back():
    cases = batch_uneven_cuboids(n=8, seed=2, n_steps=2)
    eng = _engine(cases, "mixed", (0, 512, 153600))
    sa, ma, _ = _run(eng, 2, False)
    sb, mb, lb = _run(eng, 2, True)
    assert lb["requested"] and not lb["active"]
    _same(sa, ma, sb, mb)
    eng.close()


def test_barrier_timeout_is_reported_and_the_per_pass_form_recovers():
    # a zero bound makes the first barrier that has to wait time out: every wait in the
    # kernel then ends at once, the kernel finishes, and the host reports the failure
    from spectral_galerkin_heat_amd._lib import SGHError
    cases = batch_uneven_cuboids(n=6, seed=11, n_steps=4)
    eng = _engine(cases, "mixed", (16, 512, 153600))
    sa, ma, _ = _run(eng, 2, False)
    eng.rewind()
    eng.set_loop(True)
    eng.set_loop_timeout(0.0)
    eng.run(2)
    eng.synchronize()
    assert eng.loop_state["active"]
    with pytest.raises(SGHError, match="timed out"):
        eng.metrics()
    # the error word is cleared once reported; the default bound gives the right iterates
    eng.set_loop_timeout(0.02)
    sb, mb, _ = _run(eng, 2, True)
    _same(sa, ma, sb, mb)
    eng.close()

"""Ragged and extreme shapes through both device paths, FP64 against the oracle.

Mode counts that are not multiples of the 16-wide MFMA tiles or of the 16-byte DMA unit
(33, 17, 5, 20, 36), a last row item holding a single surface row (ny = 17), the widest
surface the batched engine takes (256 × 256), two-mode axes, and a mix of all of them in
one batch. Each cuboid: relative L2 ≤ 1e-9 against the oracle with identical Picard
counts (FP64), ≤ 1e-4 for the mixed engine.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.sgh_oracle import OracleSolver, modal_rel_l2  # noqa: E402
from spectral_galerkin_heat_amd.core.laser import SingleTrackLaser  # noqa: E402
from spectral_galerkin_heat_amd.workloads import MATERIAL_316L, _cfg, make_context  # noqa: E402

SHAPES = [(33, 17, 5), (20, 36, 7), (2, 3, 2), (256, 256, 8), (17, 256, 3), (256, 4, 4),
          (144, 20, 4), (200, 33, 3), (130, 70, 3)]
# surfaces wider than 128 are split into column blocks (256: 4 × 64; 144: 3 × 48; 200:
# 64, 64, 64, 8; 130: 48, 48, 34)
SPLIT = [(144, 20, 4), (200, 33, 3), (256, 48, 4), (130, 70, 3), (64, 40, 5)]


def case(mesh, n_steps, dtype="float64"):
    size = (1.0e-3 * mesh[0] / 32 + 0.2e-3, 0.6e-3 * mesh[1] / 32 + 0.2e-3, 0.3e-3)
    cfg = _cfg(size, mesh, 5e-6, n_steps, MATERIAL_316L, 40e-6, 0.35, 300.0, 1e-8, 20, dtype)
    return cfg, SingleTrackLaser(0.2 * size[0], 0.8 * size[0], 0.5 * size[1], 1.0, 300.0)


def oracle(cfg, lp, n):
    ref = OracleSolver(make_context(cfg, lp)).initialize()
    dt = ref.ctx.num.dt
    its = [ref.step(k * dt, dt)[1]["n_evap_iter"] for k in range(n)]
    return ref, its


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_engine_ragged_and_widest_shapes(precision):
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    n = 4
    cases = [case(m, n) for m in SHAPES]
    eng = BatchEngine([make_context(c, lp) for c, lp in cases], precision=precision)
    eng.run(n)
    eng.synchronize()
    met = eng.metrics()
    for i, (cfg, lp) in enumerate(cases):
        ref, its = oracle(cfg, lp, n)
        a = eng.state(i).double().cpu().numpy()
        if precision == "fp64":
            assert modal_rel_l2(a, ref.a) < 1e-9, SHAPES[i]
            assert met[:, i, 2].astype(int).tolist() == its, SHAPES[i]
        else:
            assert modal_rel_l2(a, ref.a) < 1e-4, SHAPES[i]
    eng.close()


@pytest.mark.parametrize("variant", [(16, 256, 153600), (16, 1024, 153600), (32, 512, 153600),
                                     (0, 1024, 162816), (16, 512, 153600)])
def test_column_split_items_every_variant_fp64(variant):
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    n = 3
    cases = [case(m, n) for m in SPLIT]
    eng = BatchEngine([make_context(c, lp) for c, lp in cases], precision="fp64",
                      variant=variant)
    eng.run(n)
    eng.synchronize()
    met = eng.metrics()
    for i, (cfg, lp) in enumerate(cases):
        ref = OracleSolver(make_context(cfg, lp)).initialize()
        dt = ref.ctx.num.dt
        mets = [ref.step(k * dt, dt)[1] for k in range(n)]
        a = eng.state(i).double().cpu().numpy()
        assert modal_rel_l2(a, ref.a) < 1e-9, (SPLIT[i], variant)
        assert met[:, i, 2].astype(int).tolist() == [m["n_evap_iter"] for m in mets]
        # P_laser and T_surface_max are summed / maxed over every (row, column) item
        np.testing.assert_allclose(met[:, i, 0], [m["P_laser"] for m in mets], rtol=1e-12)
        np.testing.assert_allclose(met[:, i, 1], [m["T_surface_max"] for m in mets], rtol=1e-9)
    eng.close()


@pytest.mark.parametrize("mesh", [(33, 17, 5), (20, 36, 7), (130, 70, 3)])
def test_plain_path_ragged_shapes(mesh):
    from spectral_galerkin_heat_amd.backends import get_backend
    from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
    cfg, lp = case(mesh, 3)
    ctx = make_context(cfg, lp)
    s = SpectralSolver(get_backend("mi355x"), ctx)
    s.initialize()
    ref = OracleSolver(make_context(cfg, lp)).initialize()
    dt = ctx.num.dt
    for k in range(3):
        _, m = s.step(k * dt, dt)
        _, mr = ref.step(k * dt, dt)
        assert m["n_evap_iter"] == mr["n_evap_iter"]
    assert modal_rel_l2(s.state.a.cpu().numpy(), ref.a) < 1e-9


def test_batch_rejects_too_wide_surfaces():
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    cfg, lp = case((257, 8, 4), 2)
    with pytest.raises(ValueError, match="outside 2..256"):
        BatchEngine([make_context(cfg, lp)], precision="fp64")

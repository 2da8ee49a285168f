"""Plain-path cuboids stepped together (engine.plain_batch.PlainBatch), one stream each.

Each cuboid's kernels and their order are those of the solver run alone, so every state
and every per-step metric must equal the solo run's bit for bit: grid-mode latent heat with
T-dependent properties and top/bottom convection, box-mode latent heat with bottom
convection, a constant-property evaporation track, in FP64 and FP32.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spectral_galerkin_heat_amd.workloads import (make_context, single_track_box,  # noqa: E402
                                                  single_track_constant, single_track_tdep,
                                                  tdep_track_box)


def _cases():
    return [
        single_track_tdep(mesh=(32, 16, 8), n_steps=6),
        single_track_tdep(mesh=(48, 32, 16), n_steps=6, dtype="float32"),
        single_track_box(mesh=(40, 24, 60), n_steps=6, box_size=(0.3e-3, 0.3e-3, 0.05e-3)),
        tdep_track_box(mesh=(50, 30, 48), n_steps=6),
        single_track_constant(mesh=(64, 32, 16), n_steps=6),
    ]


def _solo(cfg, lp, n):
    from spectral_galerkin_heat_amd.backends import get_backend
    from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
    ctx = make_context(cfg, lp)
    s = SpectralSolver(get_backend("mi355x"), ctx)
    s.initialize()
    dt = ctx.num.dt
    rows = []
    for k in range(n):
        _, m = s.step(k * dt, dt)
        rows.append((float(m["P_laser"]), float(m["T_surface_max"]), float(m["n_evap_iter"]),
                     float(bool(m["picard_converged"])), float(m["picard_rel_err"])))
    torch.cuda.synchronize()
    return s.state.a.clone(), np.array(rows)


def test_plain_batch_equals_solo_runs_bit_for_bit():
    from spectral_galerkin_heat_amd.engine.plain_batch import PlainBatch
    cases = _cases()
    n = 6
    batch = PlainBatch([make_context(c, lp) for c, lp in cases])
    batch.run(4)
    batch.run(n)          # the remaining 2 steps: each cuboid stops at its own count
    batch.synchronize()
    assert batch.steps_done == [n] * len(cases)
    for i, (cfg, lp) in enumerate(cases):
        a, met = _solo(cfg, lp, n)
        b = batch.state(i)
        assert torch.equal(a, b), (i, float((a.double() - b.double()).abs().max()))
        np.testing.assert_array_equal(batch.metrics_array(i), met)
        print(f"cuboid {i} {tuple(b.shape)} {b.dtype}: identical to the solo run; "
              f"Picard counts {met[:, 2].astype(int).tolist()}")

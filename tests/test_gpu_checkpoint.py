"""Full nonlinear checkpoints and the YAML runner on the device.

A run interrupted after k steps and resumed from ``checkpoint()`` / ``save_checkpoint``
must repeat the uninterrupted run bit for bit (the checkpoint carries the modes, the
previous evaporation flux, the previous full-volume temperature, the replayed latent
modes and the box history). The runner must give the same modes as driving
``SpectralSolver`` directly over the reference's loop (``n_steps + 1`` steps at
``t = step·dt``, sgh/runner.py:395-405).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spectral_galerkin_heat_amd.workloads import (make_context, single_track_box,  # noqa: E402
                                                  single_track_constant, single_track_tdep)

CASES = {
    "constant": lambda: single_track_constant(mesh=(32, 16, 8), n_steps=8),
    "grid_latent_tdep": lambda: single_track_tdep(mesh=(48, 24, 16), n_steps=8),
    "box": lambda: single_track_box(mesh=(40, 24, 60), box_size=(0.3e-3, 0.3e-3, 0.05e-3),
                                    n_steps=8),
}


def _solver(case):
    from spectral_galerkin_heat_amd.backends import get_backend
    from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
    cfg, lp = CASES[case]()
    ctx = make_context(cfg, lp)
    s = SpectralSolver(get_backend("mi355x"), ctx)
    s.initialize()
    return s, ctx


def _advance(s, ctx, first, last):
    dt = ctx.num.dt
    m = None
    for n in range(first, last):
        _, m = s.step(n * dt, dt)
    return {k: (float(v) if k != "picard_converged" else bool(v)) for k, v in m.items()}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("via_file", [False, True])
def test_resumed_run_repeats_the_uninterrupted_run(case, via_file, tmp_path):
    s, ctx = _solver(case)
    m_ref = _advance(s, ctx, 0, 6)
    a_ref = s.state.a.cpu().numpy().copy()

    s1, ctx1 = _solver(case)
    _advance(s1, ctx1, 0, 3)
    if via_file:
        path = tmp_path / "ck.npz"
        s1.save_checkpoint(path, step=2)
    else:
        ck = s1.checkpoint(step=2)
    del s1
    s2, ctx2 = _solver(case)
    first = (s2.load_checkpoint(path) if via_file else s2.restore(ck)) + 1
    assert first == 3
    m = _advance(s2, ctx2, first, 6)
    np.testing.assert_array_equal(s2.state.a.cpu().numpy(), a_ref)
    assert m == m_ref


def test_restore_rejects_a_different_grid():
    s, _ = _solver("constant")
    ck = s.checkpoint()
    other, _ = _solver("grid_latent_tdep")
    with pytest.raises(ValueError, match="does not match"):
        other.restore(ck)


GCODE = """G21
G90
G0 X0.3 Y0.5
M3 S180
G1 X1.2 Y0.5 F30000
M5
"""


def _yaml(tmp_path, n_steps_duration=6):
    (tmp_path / "paths").mkdir()
    (tmp_path / "paths" / "track.gcode").write_text(GCODE)
    cfg = f"""
simulation:
  backend: "cpu"
  dtype: "float64"
  duration: {{value: {n_steps_duration * 1e-5}, unit: "s"}}
  dt: {{value: 1.0e-5, unit: "s"}}
  max_picard_iter: 20
  picard_tol: 1.0e-8
domain:
  size: {{value: [2.0e-3, 1.0e-3, 0.5e-3], unit: "m"}}
  mesh: {{value: [32, 16, 8], unit: ""}}
material:
  name: "316L"
  rho: 7900.0
  k: 20.0
  Cp: 500.0
  T0: 293.0
  T_liquidus: 1723.0
  T_solidus: 1658.0
  DeltaH_LV: 7.45e+6
  R_v: 149.0
  Pa: 101325.0
  T_boil: 3090.0
laser:
  radius: {{value: 50.0e-6, unit: "m"}}
  absorptivity: 0.35
  power_nominal: {{value: 180.0, unit: "W"}}
  path:
    type: "gcode"
    file: "track.gcode"
io: {{}}
"""
    p = tmp_path / "run.yaml"
    p.write_text(cfg)
    return p


def test_runner_matches_direct_solver_and_resumes(tmp_path):
    from spectral_galerkin_heat_amd.backends import get_backend
    from spectral_galerkin_heat_amd.runner import load_context, main
    from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
    cfg = _yaml(tmp_path)
    out = tmp_path / "out"
    summary = main([str(cfg), "--out", str(out), "--steps", "5", "--checkpoint-every", "3"])
    assert summary["steps"] == 6 and summary["last_step"] == 5
    modes = np.load(out / "modes.npz")["a"]
    ctx = load_context(cfg, backend="mi355x", n_steps=5)
    s = SpectralSolver(get_backend("mi355x"), ctx)
    s.initialize()
    for n in range(6):
        s.step(n * ctx.num.dt, ctx.num.dt)
    np.testing.assert_array_equal(modes, s.state.a.cpu().numpy())
    assert (out / "checkpoint_0000002.npz").exists()
    assert np.load(out / "temperature_nodes.npy").shape == (33, 17, 9)
    rows = (out / "metrics.csv").read_text().strip().splitlines()
    assert len(rows) == 7
    # resume from the step-2 checkpoint and finish: same final modes
    out2 = tmp_path / "out2"
    main([str(cfg), "--out", str(out2), "--steps", "5",
          "--resume", str(out / "checkpoint_0000002.npz")])
    np.testing.assert_array_equal(np.load(out2 / "modes.npz")["a"], modes)

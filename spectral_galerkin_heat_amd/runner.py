"""Standalone runner: YAML configuration in, time loop on the MI355X solver, files out.

Follows ``StandaloneHeatRunner.run`` (sgh/runner.py:343-429) and the CLI
(simulations/main.py:50-169): the configuration is the reference's YAML schema
(``SimulationContext.from_dict`` with the G-code path resolved next to the file),
``--backend`` overrides ``simulation.backend``, the loop runs ``n_steps + 1`` steps at
``t = step·dt`` (the reference's loop bound, ``runner.py:397``) and logs progress every
``update_interval`` of simulated time. Output is NumPy files instead of the reference's
HDF5/XDMF writers (h5py is not part of this package's environment): the final modal
state ``modes.npz`` in the reference's ``(nz, ny, nx)`` layout, the node-centred
temperature ``temperature_nodes.npy`` (``reconstruct_temperature_DCT``), per-step
metrics ``metrics.csv`` and full nonlinear checkpoints ``checkpoint_<step>.npz`` that
``--resume`` continues from bit for bit.

There is no CPU fallback: a device error ends the run with the error. Where the
reference's CLI retries on its CPU backend after a GPU out-of-memory error
(simulations/main.py:77-131), ``main`` removes the failed run's output directory, as the
reference does before its retry, and exits with status 2 and a message naming the state's
device footprint, so a caller can retry on the reference's NumPy backend or with a smaller
mesh.
"""

from __future__ import annotations

import argparse
import copy
import csv
import logging
import os
import shutil
import sys
import time

import numpy as np

__all__ = ["HeatRunner", "load_context", "main"]

logger = logging.getLogger(__name__)


def load_context(config_path, backend=None, n_steps=None):
    """SimulationContext of a YAML file (simulations/main.py:43-72)."""
    import yaml

    from .core.parameters import SimulationContext
    with open(config_path) as f:
        cfg = yaml.safe_load(f)
    cfg = copy.deepcopy(cfg)
    if backend is not None:
        cfg.setdefault("simulation", {})["backend"] = backend
    ctx = SimulationContext.from_dict(cfg, os.path.dirname(os.path.abspath(config_path)))
    if n_steps is not None:
        ctx.num.n_steps = int(n_steps)
        ctx.num.t_end = ctx.num.n_steps * ctx.num.dt
    return ctx


class HeatRunner:
    """The time loop of ``StandaloneHeatRunner.run`` around a ``SpectralSolver``."""

    def __init__(self, context, out_dir=None, checkpoint_every=0):
        from .solvers import build_solver
        self.context = context
        self.solver = build_solver(context)
        self.out_dir = out_dir
        self.checkpoint_every = int(checkpoint_every)

    def run(self, resume=None, quiet=False):
        import torch

        from .physics.spectral_helpers import reconstruct_temperature_DCT
        ctx, solver = self.context, self.solver
        solver.initialize(ctx)
        first = 0
        if resume is not None:
            first = solver.load_checkpoint(resume) + 1
        dt, n_steps = ctx.num.dt, ctx.num.n_steps
        out = self.out_dir
        rows = []
        if out is not None:
            os.makedirs(out, exist_ok=True)
        log_every = max(1, int(round(float(ctx.num.update_interval) / dt)))
        t_wall = time.perf_counter()
        last = first - 1
        for step in range(first, n_steps + 1):
            t = step * dt
            _, m = solver.step(t, dt)
            last = step
            rows.append((step, t + dt, float(m["T_surface_max"]), float(m["P_laser"]),
                         int(m["n_evap_iter"]), bool(m["picard_converged"]),
                         float(m["picard_rel_err"])))
            if not quiet and (step % log_every == 0 or step == n_steps):
                logger.info("step %d/%d t=%.4e s T_surface_max=%.1f K Picard %d (%s)", step + 1,
                            n_steps + 1, t + dt, rows[-1][2], rows[-1][4],
                            "converged" if rows[-1][5] else "cap")
            if out is not None and self.checkpoint_every > 0 and (step + 1) % self.checkpoint_every == 0:
                solver.save_checkpoint(os.path.join(out, f"checkpoint_{step:07d}.npz"), step)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_wall
        summary = {"steps": len(rows), "first_step": first, "last_step": last, "wall_s": wall,
                   "mode_steps_per_s": (float(np.prod(tuple(ctx.geom.n))) * len(rows) / wall
                                        if rows and wall > 0 else None)}
        if out is not None:
            a = solver.state.a
            np.savez(os.path.join(out, "modes.npz"), a=a.detach().cpu().numpy(),
                     step=np.int64(last), t=np.float64((last + 1) * dt))
            T = reconstruct_temperature_DCT(a, solver.state)
            np.save(os.path.join(out, "temperature_nodes.npy"), T.detach().cpu().numpy())
            solver.save_checkpoint(os.path.join(out, f"checkpoint_{last:07d}.npz"), last)
            mode = "a" if first > 0 and os.path.exists(os.path.join(out, "metrics.csv")) else "w"
            with open(os.path.join(out, "metrics.csv"), mode, newline="") as f:
                w = csv.writer(f)
                if mode == "w":
                    w.writerow(["step", "t", "T_surface_max", "P_laser", "n_evap_iter",
                                "picard_converged", "picard_rel_err"])
                w.writerows(rows)
        solver.finalize()
        return summary


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run a spectral_galerkin_heat YAML configuration "
                                             "on the MI355X backend.")
    ap.add_argument("config")
    ap.add_argument("--backend", default="mi355x",
                    help="overrides simulation.backend (this package provides 'mi355x')")
    ap.add_argument("--out", default=None, help="output directory")
    ap.add_argument("--steps", type=int, default=None, help="override the number of steps")
    ap.add_argument("--checkpoint-every", type=int, default=0)
    ap.add_argument("--resume", default=None, help="checkpoint_<step>.npz to continue from")
    args = ap.parse_args(argv)
    logging.basicConfig(level=logging.INFO, format="%(asctime)s %(levelname)s %(message)s")
    ctx = load_context(args.config, backend=args.backend, n_steps=args.steps)
    created = args.out is not None and not os.path.exists(args.out)
    try:
        summary = HeatRunner(ctx, out_dir=args.out, checkpoint_every=args.checkpoint_every).run(
            resume=args.resume)
    except Exception as exc:
        from ._lib import is_device_oom
        if not is_device_oom(exc):
            raise
        if created and os.path.isdir(args.out):
            shutil.rmtree(args.out, ignore_errors=True)
        logger.error("device out of memory: %s. State footprint of this configuration: %s. "
                     "The mi355x backend has no CPU fallback; rerun on the reference's numpy "
                     "backend or with a smaller mesh.", exc, _footprint(ctx))
        sys.exit(2)
    logger.info("done: %s", summary)
    return summary


def _footprint(ctx) -> str:
    """Rough device bytes of the plain solver's state (modal arrays and buffers)."""
    n = int(np.prod(tuple(int(v) for v in ctx.geom.n)))
    word = 8 if str(getattr(ctx.num, "dtype", "float64")).endswith("64") else 4
    return f"about {12 * n * word / 2**30:.2f} GiB for {n} modes ({word}-byte words)"


if __name__ == "__main__":
    main()

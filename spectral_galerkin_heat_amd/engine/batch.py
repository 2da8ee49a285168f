"""Batched engine: many independent cuboids advanced together on one MI355X.

``BatchEngine`` takes a list of ``SimulationContext`` objects (each with its own mode
counts, extents, material, beam, laser path and Picard controls) and runs them through
the surface-collapsed HIP engine (``sgh_batch_*`` in include/sgh_mi355x.h). Per cuboid
the step follows ``SpectralSolver.step`` (sgh/solvers/spectral.py:219-543) for the
constant-property, evaporation-only case; the whole step is captured once as a hipGraph
and replayed. Two precisions:

* ``"fp64"``  -- state, transforms and forcing in FP64 (the plain-path parity target);
* ``"mixed"`` -- state, surface transforms and forcing in FP32; E/KK regeneration,
  the Picard norms and every reduction stay in FP64.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from .. import _lib
from .._lib import CuboidSpec, call
from ..core.laser import build_laser_profile

__all__ = ["BatchEngine", "check_batchable", "laser_table", "smoke_check"]

METRIC_KEYS = ("P_laser", "T_surface_max", "n_evap_iter", "picard_converged", "picard_rel_err")


def check_batchable(ctx) -> None:
    """Raise ``ValueError`` unless the context's forcing is top-surface only."""
    mat = ctx.mat
    model = getattr(mat, "model", None)
    problems = []
    if model is not None and not model.is_constant:
        problems.append("temperature-dependent properties")
    if float(getattr(mat, "L_f", 0.0)) > 0.0:
        problems.append("latent heat (L_f > 0)")
    if float(mat.h_conv_top) > 0.0 or float(mat.h_conv_bottom) > 0.0:
        problems.append("convection")
    if ctx.fine is not None:
        problems.append("a fine_mesh box")
    if int(ctx.geom.n.x) > 256 or int(ctx.geom.n.y) > 256 or int(ctx.geom.n.x) < 2:
        problems.append("surface mode counts outside 2..256")
    if problems:
        raise ValueError("the batched engine covers top-surface forcing only; this case has "
                         + ", ".join(problems) + " (use SpectralSolver on the mi355x backend)")


def laser_table(ctx, n_steps=None):
    """Per-step beam rows {x, y, power, on, vx, vy} from ``ctx.laser_path`` at t = n·dt."""
    n_steps = int(ctx.num.n_steps if n_steps is None else n_steps)
    dt = float(ctx.num.dt)
    rows = np.zeros((n_steps, 6), dtype=np.float64)
    for n in range(n_steps):
        ls = ctx.laser_path.get_state(n * dt, dt)
        rows[n] = (float(ls.x), float(ls.y), float(ls.power), 1.0 if ls.is_on else 0.0,
                   float(ls.v[0]), float(ls.v[1]))
    return rows


class BatchEngine:
    def __init__(self, contexts, precision: str = "mixed", n_steps=None, metrics_rows=None,
                 omega=None, tol=None, max_it=None):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if not torch.cuda.is_available():
            raise RuntimeError("BatchEngine needs the MI355X (torch.cuda is unavailable)")
        _lib.lib.load()
        self.contexts = list(contexts)
        self.n = len(self.contexts)
        if self.n == 0:
            raise ValueError("empty batch")
        self.precision = precision
        self.code = _lib.F64 if precision == "fp64" else _lib.F32
        self.tdtype = torch.float64 if precision == "fp64" else torch.float32
        self.stream = torch.cuda.Stream()
        steps = [int(c.num.n_steps if n_steps is None else n_steps) for c in self.contexts]
        self.n_steps = steps
        self.metrics_rows = int(max(steps) if metrics_rows is None else metrics_rows)
        self._tables = []
        specs = (CuboidSpec * self.n)()
        self.shapes = []
        for i, ctx in enumerate(self.contexts):
            check_batchable(ctx)
            num, geom, mat, lp = ctx.num, ctx.geom, ctx.mat, ctx.laser
            prof = build_laser_profile(lp.profile, lp.super_gaussian_order,
                                       cell_integrated=lp.cell_integrated)
            tab = np.ascontiguousarray(laser_table(ctx, steps[i]))
            self._tables.append(tab)
            sp = specs[i]
            sp.nx, sp.ny, sp.nz = int(geom.n.x), int(geom.n.y), int(geom.n.z)
            self.shapes.append((sp.nz, sp.ny, sp.nx))
            sp.max_it = int(max_it if max_it is not None else
                            (num.max_picard_iter if num.max_picard_iter is not None else 30))
            sp.n_steps = steps[i]
            sp.cell_integrated = int(bool(prof.cell_integrated))
            sp.Lx, sp.Ly, sp.Lz = (float(v) for v in geom.size)
            sp.k, sp.rho, sp.cp = float(mat.k), float(mat.rho), float(mat.Cp)
            sp.Pa, sp.T_boil, sp.dH = float(mat.Pa), float(mat.T_boil), float(mat.DeltaH_LV)
            sp.R_v, sp.T_liq = float(mat.R_v), float(mat.T_liquidus)
            sp.absorptivity, sp.rx, sp.ry = float(lp.absorptivity), float(lp.r_x), float(lp.r_y)
            sp.order = float(prof.order)
            sp.dt = float(num.dt)
            sp.omega = float(omega if omega is not None else
                             (num.picard_omega if num.picard_omega is not None else 0.1))
            sp.tol = float(tol if tol is not None else
                           (num.picard_tol if num.picard_tol is not None else 1e-4))
            if self.code == _lib.F32:
                # the reference compares against dtype(tol)
                sp.tol = float(np.float32(sp.tol))
            sp.laser = tab.ctypes.data_as(C.POINTER(C.c_double))
        self._specs = specs
        plan = C.c_void_p()
        call("sgh_batch_create", self.code, self.n, C.cast(specs, C.c_void_p), self.metrics_rows,
             C.byref(plan), self.stream.cuda_stream)
        self._plan = plan
        info = (C.c_longlong * 6)()
        call("sgh_batch_info", self._plan, C.cast(info, C.c_void_p))
        self.info = {"volume_elems": info[0], "surface_elems": info[1], "surface_items": info[2],
                     "volume_items": info[3], "max_it": info[4], "lds_bytes": info[5]}
        self.mode_counts = [int(np.prod(s)) for s in self.shapes]
        self.steps_done = 0
        self.reset_initial_condition()

    # -- state ------------------------------------------------------------------------
    def reset_initial_condition(self):
        """a = 0 except a[0,0,0] = T0·√(LxLyLz) (sgh/solvers/spectral.py:176-180)."""
        with torch.cuda.stream(self.stream):
            for i, ctx in enumerate(self.contexts):
                a = torch.zeros(self.shapes[i], dtype=self.tdtype, device="cuda")
                g = ctx.geom
                a[0, 0, 0] = float(ctx.mat.T0) * math.sqrt(float(g.size.x) * float(g.size.y) * float(g.size.z))
                self.set_state(i, a)
        self.stream.synchronize()

    def set_state(self, i, a):
        a = a.to(device="cuda", dtype=self.tdtype).contiguous()
        if tuple(a.shape) != self.shapes[i]:
            raise ValueError(f"state of cuboid {i} must have shape {self.shapes[i]}")
        self.stream.wait_stream(torch.cuda.current_stream())
        call("sgh_batch_set_state", self._plan, i, a.data_ptr(), self.stream.cuda_stream)
        a.record_stream(self.stream)

    def state(self, i):
        with torch.cuda.stream(self.stream):
            out = torch.empty(self.shapes[i], dtype=self.tdtype, device="cuda")
            call("sgh_batch_get_state", self._plan, i, out.data_ptr(), self.stream.cuda_stream)
        torch.cuda.current_stream().wait_stream(self.stream)
        return out

    # -- stepping ---------------------------------------------------------------------
    def run(self, n_steps: int, use_graph: bool = True):
        """Enqueue ``n_steps`` steps for every live cuboid (asynchronous)."""
        self.stream.wait_stream(torch.cuda.current_stream())
        call("sgh_batch_run", self._plan, int(n_steps), int(bool(use_graph)), self.stream.cuda_stream)
        self.steps_done += int(n_steps)

    def synchronize(self):
        self.stream.synchronize()

    def metrics(self, rows=None):
        """Per-step, per-cuboid metrics array [rows, n, 5] (P_laser, T_surface_max,
        n_evap_iter, picard_converged, picard_rel_err)."""
        rows = min(self.steps_done, self.metrics_rows) if rows is None else int(rows)
        out = np.zeros((rows, self.n, 5), dtype=np.float64)
        if rows:
            call("sgh_batch_read_metrics", self._plan, out.ctypes.data, rows, self.stream.cuda_stream)
        return out

    def set_timing(self, enable: bool):
        call("sgh_batch_set_timing", self._plan, int(bool(enable)))

    def timing(self):
        out = (C.c_double * 4)()
        call("sgh_batch_timing", self._plan, C.cast(out, C.c_void_p))
        return {"surface_ms": out[0], "surface_launches": int(out[1]),
                "volume_ms": out[2], "volume_launches": int(out[3])}

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            self.stream.synchronize()
            call("sgh_batch_destroy", self._plan)
            self._plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def smoke_check():
    """Three uneven cuboids, FP64 engine vs the oracle for 3 steps (rel-L2 ≤ 1e-9)."""
    from oracle.sgh_oracle import OracleSolver
    from ..workloads import batch_uneven_cuboids, make_context
    cases = batch_uneven_cuboids(n=3, seed=11, n_steps=3)
    ctxs = [make_context(c, lp) for c, lp in cases]
    eng = BatchEngine(ctxs, precision="fp64")
    eng.run(3)
    eng.synchronize()
    worst = 0.0
    for i, (c, lp) in enumerate(cases):
        ref = OracleSolver(make_context(c, lp)).initialize()
        for n in range(3):
            ref.step(n * ref.ctx.num.dt, ref.ctx.num.dt)
        a = eng.state(i).double().cpu().numpy()
        worst = max(worst, float(np.linalg.norm(a - ref.a) / np.linalg.norm(ref.a)))
    eng.close()
    print(f"smoke: batched FP64 engine, 3 cuboids x 3 steps, worst rel-L2 vs oracle = {worst:.3e}",
          flush=True)
    assert worst < 1e-9, worst

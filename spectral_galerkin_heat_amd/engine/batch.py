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

__all__ = ["BatchEngine", "BatchPool", "PoolFull", "check_batchable", "footprint_bytes",
           "laser_table", "smoke_check"]

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


def cuboid_spec(ctx, n_steps, precision_code, omega=None, tol=None, max_it=None):
    """``sgh_cuboid_spec`` and its laser table (the table must outlive the spec's use)."""
    check_batchable(ctx)
    num, geom, mat, lp = ctx.num, ctx.geom, ctx.mat, ctx.laser
    prof = build_laser_profile(lp.profile, lp.super_gaussian_order,
                               cell_integrated=lp.cell_integrated)
    tab = np.ascontiguousarray(laser_table(ctx, n_steps))
    sp = CuboidSpec()
    sp.nx, sp.ny, sp.nz = int(geom.n.x), int(geom.n.y), int(geom.n.z)
    sp.max_it = int(max_it if max_it is not None else
                    (num.max_picard_iter if num.max_picard_iter is not None else 30))
    sp.n_steps = int(n_steps)
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
    if precision_code == _lib.F32:
        sp.tol = float(np.float32(sp.tol))  # the reference compares against dtype(tol)
    sp.laser = tab.ctypes.data_as(C.POINTER(C.c_double))
    return sp, tab


def initial_modes(ctx, shape, tdtype):
    """a = 0 except a[0,0,0] = T0·√(LxLyLz) (sgh/solvers/spectral.py:176-180)."""
    a = torch.zeros(shape, dtype=tdtype, device="cuda")
    g = ctx.geom
    a[0, 0, 0] = float(ctx.mat.T0) * math.sqrt(float(g.size.x) * float(g.size.y) * float(g.size.z))
    return a


def footprint_bytes(ctx, n_steps, precision="mixed"):
    """Device bytes a cuboid takes in a pool: state, 10 surface fields, laser table."""
    w = 8 if precision == "fp64" else 4
    nx, ny, nz = (int(v) for v in ctx.geom.n)
    return w * (nx * ny * nz + 10 * nx * ny) + 48 * max(1, int(n_steps))


class BatchEngine:
    """A batch of cuboids advanced together; fixed membership unless built as a pool."""

    def __init__(self, contexts, precision: str = "mixed", n_steps=None, metrics_rows=None,
                 omega=None, tol=None, max_it=None, variant=None):
        """``variant``: None keeps the default surface-transform variant, ``"auto"`` runs
        the autotuner on a scratch copy of this batch, or a ``(rows, threads, lds_bytes)``
        tuple / dict sets one explicitly."""
        self._setup(precision)
        self.contexts = list(contexts)
        self.n = len(self.contexts)
        if self.n == 0:
            raise ValueError("empty batch")
        steps = [int(c.num.n_steps if n_steps is None else n_steps) for c in self.contexts]
        self.n_steps = steps
        self.metrics_rows = int(max(steps) if metrics_rows is None else metrics_rows)
        specs = (CuboidSpec * self.n)()
        tables = []
        self.shapes = []
        for i, ctx in enumerate(self.contexts):
            sp, tab = cuboid_spec(ctx, steps[i], self.code, omega, tol, max_it)
            tables.append(tab)
            specs[i] = sp
            self.shapes.append((sp.nz, sp.ny, sp.nx))
        plan = C.c_void_p()
        call("sgh_batch_create", self.code, self.n, C.cast(specs, C.c_void_p), self.metrics_rows,
             C.byref(plan), self.stream.cuda_stream)
        self._plan = plan
        self.mode_counts = [int(np.prod(s)) for s in self.shapes]
        self.steps_done = 0
        self.tuning = None
        loop = False
        if isinstance(variant, str) and variant == "auto":
            from .autotune import autotune
            self.tuning = autotune(self.contexts, precision)
            variant, loop = self.tuning.variant, self.tuning.loop
        if variant is not None:
            if isinstance(variant, dict):
                loop = bool(variant.get("loop", False))
                variant = (variant["rows"], variant["threads"], variant["lds_bytes"])
            self.set_variant(*variant)
        if loop:
            self.set_loop(True)
        self.reset_initial_condition()

    def _setup(self, precision):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if not torch.cuda.is_available():
            raise RuntimeError("BatchEngine needs the MI355X (torch.cuda is unavailable)")
        _lib.lib.load()
        self.precision = precision
        self.code = _lib.F64 if precision == "fp64" else _lib.F32
        self.tdtype = torch.float64 if precision == "fp64" else torch.float32
        # graph capture needs a non-default stream; default priority
        self.stream = torch.cuda.Stream()
        self._plan = C.c_void_p()

    @property
    def info(self):
        info = (C.c_longlong * 6)()
        call("sgh_batch_info", self._plan, C.cast(info, C.c_void_p))
        return {"live_state_elems": info[0], "surface_capacity": info[1],
                "surface_items": info[2], "volume_items": info[3], "max_it": info[4],
                "lds_bytes": info[5]}

    # -- surface-transform variant -----------------------------------------------------
    def set_variant(self, rows, threads, lds_bytes=150 * 1024):
        """Rows per transform item, threads per workgroup and LDS bytes per workgroup of
        the batched surface transforms (takes effect at the next ``run``). ``rows = -1``
        selects the resident step kernel (one workgroup per cuboid runs its whole Picard
        loop; mixed precision, surfaces up to 128 × 128, otherwise the per-pass launches of
        16-row items with the given LDS budget run)."""
        call("sgh_batch_set_variant", self._plan, int(rows), int(threads), int(lds_bytes),
             self.stream.cuda_stream)

    @property
    def variant(self):
        out = (C.c_int * 3)()
        call("sgh_batch_get_variant", self._plan, C.cast(out, C.c_void_p))
        return {"rows": out[0], "threads": out[1], "lds_bytes": out[2]}

    def set_loop(self, enable: bool):
        """Run each step's surface work as one persistent launch (the items of a cuboid
        meet at device-scope barriers) instead of one launch per pass. Used only while the
        work list fits one workgroup per CU and for the 16-row variants with 512 or 1024
        threads; otherwise the per-pass launches run. Same iterates either way."""
        call("sgh_batch_set_loop", self._plan, int(bool(enable)), self.stream.cuda_stream)

    def set_loop_timeout(self, seconds: float):
        """Bound of one barrier wait of the persistent kernel (default 0.02 s). A wait that
        times out is reported by the next ``steps``/``metrics`` read as an error."""
        call("sgh_batch_set_loop_timeout", self._plan, float(seconds), self.stream.cuda_stream)

    @property
    def loop_state(self):
        """{'requested', 'active'}; 'active' reflects the work lists of the last run."""
        out = (C.c_int * 2)()
        call("sgh_batch_loop_state", self._plan, C.cast(out, C.c_void_p))
        return {"requested": bool(out[0]), "active": bool(out[1])}

    def set_team_max(self, team_max: int):
        """Largest team of workgroups a cuboid's step may run on in the resident step kernel
        (1 to 8; default 8). Team sizes are chosen per work list; results are bit-identical
        for every team size."""
        call("sgh_batch_set_team_max", self._plan, int(team_max), self.stream.cuda_stream)

    @property
    def team_sizes(self):
        """Team size of each slot in the last work lists (0: free, or stepped by the
        per-pass launches)."""
        n = int(call("sgh_batch_team_sizes", self._plan, None, 0))
        out = (C.c_int * max(n, 1))()
        call("sgh_batch_team_sizes", self._plan, C.cast(out, C.c_void_p), n)
        return [int(out[i]) for i in range(n)]

    def rewind(self):
        """Restart every cuboid from its initial condition at step 0."""
        call("sgh_batch_rewind", self._plan, self.stream.cuda_stream)
        self.steps_done = 0
        self.reset_initial_condition()

    # -- state ------------------------------------------------------------------------
    def reset_initial_condition(self):
        with torch.cuda.stream(self.stream):
            for i, ctx in enumerate(self.contexts):
                self.set_state(i, initial_modes(ctx, self.shapes[i], self.tdtype))
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

    def cuboid_metrics(self, i, rows):
        out = np.zeros((int(rows), 5), dtype=np.float64)
        if rows:
            call("sgh_batch_read_metrics", self._plan, int(i), out.ctypes.data, int(rows),
                 self.stream.cuda_stream)
        return out

    def metrics(self, rows=None):
        """Per-step, per-cuboid metrics array [rows, n, 5] (P_laser, T_surface_max,
        n_evap_iter, picard_converged, picard_rel_err)."""
        rows = min(self.steps_done, self.metrics_rows) if rows is None else int(rows)
        out = np.zeros((rows, self.n, 5), dtype=np.float64)
        for i in range(self.n):
            out[:, i] = self.cuboid_metrics(i, rows)
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


class BatchPool(BatchEngine):
    """Continuous batching: cuboids are added to free slots and removed when done.

    Capacities are in elements (state, surface) and laser-table rows; ``add`` raises
    ``PoolFull`` when the pool has no room for a request.
    """

    def __init__(self, precision="mixed", max_cuboids=128, state_capacity=1 << 25,
                 surface_capacity=1 << 21, laser_capacity=1 << 18, max_nz=64, metrics_rows=2048):
        self._setup(precision)
        self.metrics_rows = int(metrics_rows)
        self.max_cuboids = int(max_cuboids)
        self.max_nz = int(max_nz)
        plan = C.c_void_p()
        call("sgh_batch_pool_create", self.code, self.max_cuboids, int(state_capacity),
             int(surface_capacity), int(laser_capacity), self.max_nz, self.metrics_rows,
             C.byref(plan))
        self._plan = plan
        self.contexts, self.shapes, self.n_steps = {}, {}, {}
        self.steps_done = 0

    def check_admissible(self, ctx, n_steps):
        """Raise ``ValueError`` for a request this pool can never hold, whatever is freed:
        a cuboid deeper than ``max_nz``, more steps than the metrics rows, or forcing the
        batched engine does not cover."""
        check_batchable(ctx)
        nz = int(ctx.geom.n.z)
        if nz > self.max_nz:
            raise ValueError(f"nz = {nz} exceeds the pool's max_nz = {self.max_nz}")
        if int(n_steps) > self.metrics_rows:
            raise ValueError(f"{int(n_steps)} steps exceed the pool's {self.metrics_rows} "
                             "metrics rows")

    def matrix_sizes(self):
        """Distinct surface mode counts whose cosine matrices the pool holds now."""
        n = call("sgh_batch_matrix_sizes", self._plan, None, 0)
        out = (C.c_int * max(1, n))()
        n = call("sgh_batch_matrix_sizes", self._plan, C.cast(out, C.c_void_p), n)
        return sorted(out[:n])

    def add(self, ctx, n_steps, initial_state=None):
        self.check_admissible(ctx, n_steps)
        sp, tab = cuboid_spec(ctx, n_steps, self.code)
        slot = C.c_int(-1)
        rc = _lib.lib.sgh_batch_add(self._plan, C.byref(sp), C.byref(slot), self.stream.cuda_stream)
        if rc == -4:
            raise PoolFull(_lib.last_error())
        if rc < 0:
            raise _lib.SGHError(f"sgh_batch_add failed ({rc}): {_lib.last_error()}")
        i = slot.value
        self.contexts[i] = ctx
        self.shapes[i] = (sp.nz, sp.ny, sp.nx)
        self.n_steps[i] = int(n_steps)
        a = initial_modes(ctx, self.shapes[i], self.tdtype) if initial_state is None else initial_state
        self.set_state(i, a)
        return i

    def rewind(self):
        raise NotImplementedError("a pool's members start when they are added")

    def remove(self, i):
        call("sgh_batch_remove", self._plan, int(i), self.stream.cuda_stream)
        self.contexts.pop(i, None)
        self.shapes.pop(i, None)
        self.n_steps.pop(i, None)

    def steps(self):
        out = (C.c_int * self.max_cuboids)()
        call("sgh_batch_steps_done", self._plan, C.cast(out, C.c_void_p), self.stream.cuda_stream)
        return list(out)


class PoolFull(RuntimeError):
    """The pool has no free slot or no contiguous room for the request."""

"""Many plain-path solvers advanced together: the physics the surface-collapsed engine does
not cover (grid-mode and box-mode latent heat, top and bottom convection,
temperature-dependent properties with the property correction) for independent cuboids.

Each cuboid is an unchanged ``SpectralSolver`` on the mi355x backend
(``sgh/solvers/spectral.py:219-543`` through ``kernels.picard_step``) bound to a HIP stream
of its own. One host thread round-robins over the cuboids' steps
(``kernels.picard_step_steps``, the plain step as a generator): a cuboid whose next Picard
decision waits for a residual the device has not written yet gives the host to the next
one, so the small kernels of all cuboids' iterations are in flight on the device at once
and fill CUs that one cuboid's launches leave idle. A cuboid's kernels and their order on
its stream are those of the solver run alone, so its results are the same bit for bit
(tested). The reference has no batched engine; this has no counterpart there.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from ..backends import get_backend
from ..physics.mi355x_kernels import picard_step_steps
from ..solvers.spectral import SpectralSolver

__all__ = ["PlainBatch", "PlainLane"]


class PlainBatch:
    """Independent plain-path cuboids stepped together, one stream each."""

    def __init__(self, contexts, n_steps=None):
        if not torch.cuda.is_available():
            raise RuntimeError("PlainBatch needs the MI355X (torch.cuda is unavailable)")
        self.contexts = list(contexts)
        if not self.contexts:
            raise ValueError("empty batch")
        backend = get_backend("mi355x")
        self.solvers = []
        for ctx in self.contexts:
            s = SpectralSolver(backend, ctx)
            s._plain_fork = False   # concurrent solvers: no side stream per solver
            s.initialize()
            self.solvers.append(s)
        # default priority (the only one used anywhere in this package)
        self.streams = [torch.cuda.Stream() for _ in self.solvers]
        # every solver's tables were built on the current stream: the solver streams start
        # after them
        cur = torch.cuda.current_stream()
        for st in self.streams:
            st.wait_stream(cur)
        self.n_steps = [int(c.num.n_steps if n_steps is None else n_steps)
                        for c in self.contexts]
        self.steps_done = [0] * len(self.solvers)
        self.metrics = [[] for _ in self.solvers]

    @property
    def n(self):
        return len(self.solvers)

    def _steps(self, i, n):
        s, ctx = self.solvers[i], self.contexts[i]
        dt = ctx.num.dt
        for _ in range(n):
            k = self.steps_done[i]
            if k >= self.n_steps[i]:
                return
            _, m = yield from picard_step_steps(s, k * dt, dt)
            self.metrics[i].append(m)
            self.steps_done[i] = k + 1

    def run(self, n_steps: int, idle_sleep_s: float = 0.0):
        """Advance every cuboid by ``n_steps`` steps (fewer where its own step count ends
        first); returns when all have enqueued them (``synchronize`` waits for the device)."""
        gens = {i: self._steps(i, int(n_steps)) for i in range(self.n)}
        while gens:
            progressed = False
            for i in list(gens):
                with torch.cuda.stream(self.streams[i]):
                    before = self.steps_done[i]
                    try:
                        next(gens[i])
                    except StopIteration:
                        del gens[i]
                        progressed = True
                        continue
                    progressed = progressed or self.steps_done[i] != before
            if not progressed and idle_sleep_s > 0:
                time.sleep(idle_sleep_s)

    def synchronize(self):
        for st in self.streams:
            st.synchronize()

    def state(self, i):
        """Modal state of cuboid i ([nz, ny, nx], the solver's dtype), after the device
        has finished its steps."""
        self.streams[i].synchronize()
        return self.solvers[i].state.a

    def metrics_array(self, i):
        """Per-step metrics of cuboid i: rows (P_laser, T_surface_max, n_evap_iter,
        picard_converged, picard_rel_err)."""
        self.streams[i].synchronize()
        rows = [(float(m["P_laser"]), float(m["T_surface_max"]), float(m["n_evap_iter"]),
                 float(bool(m["picard_converged"])), float(m["picard_rel_err"]))
                for m in self.metrics[i]]
        return np.array(rows, dtype=np.float64).reshape(-1, 5)


class PlainLane:
    """The plain-path side of the request loop: cuboids join and leave between steps.

    Each admitted request gets a solver on a stream of its own; ``advance`` moves every
    member forward by whole steps, round-robining the host over their Picard loops as
    ``PlainBatch.run`` does, and reports the requests whose last step has been enqueued.
    """

    def __init__(self, max_active=16):
        self.max_active = int(max_active)
        self._members = {}      # key -> dict(solver, ctx, stream, n_steps, done, metrics, gen)
        self._backend = None

    def __len__(self):
        return len(self._members)

    def has_room(self):
        return len(self._members) < self.max_active

    def add(self, key, ctx, n_steps):
        if self._backend is None:
            self._backend = get_backend("mi355x")
        s = SpectralSolver(self._backend, ctx)
        s._plain_fork = False   # concurrent solvers: no side stream per solver
        s.initialize()
        st = torch.cuda.Stream()
        st.wait_stream(torch.cuda.current_stream())
        m = {"solver": s, "ctx": ctx, "stream": st, "n_steps": int(n_steps), "done": 0,
             "metrics": [], "gen": None}
        self._members[key] = m

    def _one_step(self, m):
        dt = m["ctx"].num.dt
        _, met = yield from picard_step_steps(m["solver"], m["done"] * dt, dt)
        m["metrics"].append(met)
        m["done"] += 1

    def advance(self, steps=1):
        """Every member runs up to ``steps`` more steps (fewer if its count ends first);
        returns the keys of members whose last step is enqueued."""
        todo = {k: min(int(steps), m["n_steps"] - m["done"]) for k, m in self._members.items()}
        todo = {k: v for k, v in todo.items() if v > 0}
        while todo:
            for k in list(todo):
                m = self._members[k]
                with torch.cuda.stream(m["stream"]):
                    if m["gen"] is None:
                        m["gen"] = self._one_step(m)
                    try:
                        next(m["gen"])
                    except StopIteration:
                        m["gen"] = None
                        todo[k] -= 1
                        if todo[k] == 0:
                            del todo[k]
        return [k for k, m in self._members.items() if m["done"] >= m["n_steps"]]

    def take(self, key):
        """Final modal state (device) and per-step metrics [n_steps, 5] of a finished
        member, which leaves the lane."""
        m = self._members.pop(key)
        m["stream"].synchronize()
        rows = [(float(x["P_laser"]), float(x["T_surface_max"]), float(x["n_evap_iter"]),
                 float(bool(x["picard_converged"])), float(x["picard_rel_err"]))
                for x in m["metrics"]]
        return m["solver"].state.a, np.array(rows, dtype=np.float64).reshape(-1, 5)

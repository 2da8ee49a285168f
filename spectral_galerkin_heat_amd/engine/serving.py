"""Request loop: continuous batching of scan simulations on one MI355X.

Requests (a ``SimulationContext`` with its laser path and a step count) arrive over
time. The server keeps every admitted simulation's modal state resident in HBM inside a
``BatchPool`` and advances all of them together, a few steps per tick. Between ticks it

* retires simulations whose last step has run (final state and per-step metrics are
  read back; completion latency = completion time − arrival time),
* admits queued requests in arrival order while the pool has a free slot and room in
  its packed arrays (admission by memory footprint; a request that does not fit waits),

so requests that arrive while others are on the GPU join the running batch and their
work overlaps on the device. Requests whose physics the surface-collapsed pool does not
cover (latent heat, convection, temperature-dependent properties, a fine box) go to a
plain lane (``PlainLane``: an unchanged ``SpectralSolver`` per request on a stream of its
own, ``plain_slots`` at a time), which advances one step per tick beside the pool. A simulation's arithmetic depends only on its own data
(fixed per-cuboid work decomposition and reduction order), so its result is the same
alone and in any mix. Every ``check_every``-th request is re-solved alone in FP64 and
its relative L2 against the served (mixed-precision) result is recorded.
"""

from __future__ import annotations

import heapq
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .._lib import DeviceOutOfMemory
from .batch import BatchEngine, BatchPool, PoolFull, check_batchable, footprint_bytes
from .plain_batch import PlainLane

__all__ = ["Request", "ScanServer", "poisson_arrivals"]


@dataclass
class Request:
    rid: int
    ctx: object
    n_steps: int
    arrival: float                 # seconds after the server clock started
    slot: int = -1
    start_step: int = 0            # pool step counter value at admission
    admitted_at: float = float("nan")
    completed_at: float = float("nan")
    result: object = None          # final modal state (torch, device)
    metrics: object = None         # [n_steps, 5] numpy
    fp64_rel_l2: float = float("nan")
    footprint: int = 0
    plain: bool = False            # served by the plain lane

    @property
    def latency(self) -> float:
        return self.completed_at - self.arrival

    @property
    def modes(self) -> int:
        return int(np.prod(tuple(self.ctx.geom.n)))


def poisson_arrivals(rate_hz: float, n: int, rng) -> np.ndarray:
    """Arrival times of a Poisson process of the given rate (first at t = 0)."""
    gaps = rng.exponential(1.0 / rate_hz, size=n)
    gaps[0] = 0.0
    return np.cumsum(gaps)


@dataclass
class ServeReport:
    requests: list = field(default_factory=list)
    wall_s: float = 0.0
    mode_steps: int = 0
    device_steps: int = 0
    peak_active: int = 0          # most requests resident in the pool at once
    peak_bytes: int = 0           # most pool bytes (state, surface fields, laser rows) in use

    @property
    def latencies(self):
        return np.array([r.latency for r in self.requests])

    def summary(self):
        lat = self.latencies
        wait = np.array([r.admitted_at - r.arrival for r in self.requests])
        return {
            "queue_wait_p50_s": float(np.percentile(wait, 50)) if len(wait) else None,
            "queue_wait_max_s": float(wait.max()) if len(wait) else None,
            "peak_active": self.peak_active,
            "peak_pool_bytes": self.peak_bytes,
            "requests": len(self.requests),
            "plain_requests": sum(1 for r in self.requests if r.plain),
            "plain_p50_latency_s": (float(np.percentile([r.latency for r in self.requests
                                                         if r.plain], 50))
                                    if any(r.plain for r in self.requests) else None),
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_latency_s": float(lat.mean()) if len(lat) else None,
            "aggregate_mode_steps_per_s": self.mode_steps / self.wall_s if self.wall_s else None,
            "wall_s": self.wall_s,
            "device_steps": self.device_steps,
            "fp64_checks": [r.fp64_rel_l2 for r in self.requests if not np.isnan(r.fp64_rel_l2)],
        }


class ScanServer:
    """Continuous-batching request loop over a ``BatchPool`` plus a plain lane.

    ``variant``: the pool's surface-transform variant, ``(rows, threads, lds_bytes)``;
    ``"resident"`` (the default in mixed precision) runs every cuboid whose surface sides are
    at most 128 on the resident step kernel and the wider ones on 16-row per-pass launches in
    the same step; None keeps the per-pass default. A request's result depends only on its
    own cuboid and the variant, never on what else is in the pool.
    """
    MIN_BUDGET = 64 << 20   # smallest pool budget tried after device out-of-memory errors
    RESIDENT = (-1, 512, 150 * 1024)

    def __init__(self, precision="mixed", max_cuboids=128, hbm_budget_bytes=8 << 30,
                 max_nz=64, max_steps=2048, steps_per_tick=8, check_every=20,
                 keep_results=True, plain_slots=16, variant="resident"):
        word = 8 if precision == "fp64" else 4
        # split the budget between the packed state and the surface fields; a pool the
        # device cannot hold is retried at half the budget (down to MIN_BUDGET)
        budget = int(hbm_budget_bytes)
        while True:
            state_cap = int(budget * 0.85) // word
            surf_cap = int(budget * 0.15) // (10 * word)
            try:
                self.pool = BatchPool(precision=precision, max_cuboids=max_cuboids,
                                      state_capacity=state_cap, surface_capacity=surf_cap,
                                      laser_capacity=max_cuboids * max_steps, max_nz=max_nz,
                                      metrics_rows=max_steps)
                break
            except DeviceOutOfMemory:
                if budget // 2 < self.MIN_BUDGET:
                    raise
                budget //= 2
        if isinstance(variant, str) and variant == "resident":
            variant = self.RESIDENT if precision == "mixed" else None
        if variant is not None:
            self.pool.set_variant(*variant)
        self.variant = self.pool.variant
        self.hbm_budget_bytes = budget
        self.precision = precision
        self.max_steps = max_steps
        self.steps_per_tick = int(steps_per_tick)
        self.check_every = int(check_every)
        self.keep_results = keep_results
        self.device_steps = 0
        self._last_refusal = ""
        self.plain = PlainLane(plain_slots) if plain_slots > 0 else None

    def _admit(self, req: Request, now: float) -> bool:
        try:
            slot = self.pool.add(req.ctx, req.n_steps)
        except PoolFull as exc:
            self._last_refusal = str(exc)
            return False
        req.slot = slot
        req.admitted_at = now
        req.start_step = self.device_steps
        req.footprint = footprint_bytes(req.ctx, req.n_steps, self.precision)
        return True

    def serve(self, requests, clock=time.perf_counter, sleep=time.sleep) -> ServeReport:
        """Serve ``requests`` (sorted by arrival) in real time; returns the report."""
        queue = sorted(requests, key=lambda r: r.arrival)
        pending = list(queue)
        waiting: list = []
        active: dict = {}             # slot -> request
        done: list = []
        for r in pending:
            try:
                check_batchable(r.ctx)
            except ValueError:
                if self.plain is None:
                    raise
                r.plain = True
                continue
            if r.n_steps > self.max_steps:
                raise ValueError(f"request {r.rid}: {r.n_steps} steps > max_steps {self.max_steps}")
            try:
                self.pool.check_admissible(r.ctx, r.n_steps)
            except ValueError as exc:
                raise ValueError(f"request {r.rid}: {exc}") from None
        pwaiting: list = []
        pactive: dict = {}            # rid -> request (plain lane)
        t0 = clock()
        finish_heap: list = []        # (device step at which it completes, rid)
        by_rid = {r.rid: r for r in queue}
        peak_active = peak_bytes = 0
        while pending or waiting or active or pwaiting or pactive:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.pop(0)
                (pwaiting if r.plain else waiting).append(r)
            while pwaiting and self.plain.has_room():
                r = pwaiting.pop(0)
                self.plain.add(r.rid, r.ctx, r.n_steps)
                r.admitted_at = now
                r.footprint = 12 * r.modes * np.dtype(r.ctx.num.dtype).itemsize
                pactive[r.rid] = r
            admitted = []
            for r in list(waiting):
                if not self._admit(r, now):
                    break                    # FIFO: later requests wait behind this one
                waiting.remove(r)
                active[r.slot] = r
                heapq.heappush(finish_heap, (r.start_step + r.n_steps, r.rid))
                admitted.append(r)
            peak_active = max(peak_active, len(active) + len(pactive))
            peak_bytes = max(peak_bytes, sum(r.footprint for r in active.values()))
            if not active and waiting:
                raise RuntimeError(f"request {waiting[0].rid} does not fit an empty pool "
                                   f"(raise hbm_budget_bytes): {self._last_refusal}")
            if not active and not pactive:
                if pending:
                    sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            next_finish = None
            if active:
                # run until the next completion, at most steps_per_tick steps (enqueued; the
                # device runs them while the host drives the plain lane)
                next_finish = finish_heap[0][0]
                k = max(1, min(self.steps_per_tick, next_finish - self.device_steps))
                self.pool.run(k)
                self.device_steps += k
            if pactive:
                # one step of every plain request; its kernels run beside the pool's
                for rid in self.plain.advance(1):
                    r = pactive.pop(rid)
                    a, met = self.plain.take(rid)
                    r.completed_at = clock() - t0
                    if self.keep_results:
                        r.result = a.clone()
                    r.metrics = met
                    done.append(r)
            if not active:
                continue
            if self.device_steps >= next_finish or pending:
                self.pool.synchronize()
            while finish_heap and finish_heap[0][0] <= self.device_steps:
                _, rid = heapq.heappop(finish_heap)
                r = by_rid[rid]
                self.pool.synchronize()
                r.completed_at = clock() - t0
                if self.keep_results:
                    r.result = self.pool.state(r.slot).clone()
                r.metrics = self.pool.cuboid_metrics(r.slot, r.n_steps)
                self.pool.remove(r.slot)
                del active[r.slot]
                done.append(r)
        wall = clock() - t0
        rep = ServeReport(requests=sorted(done, key=lambda r: r.rid), wall_s=wall,
                          mode_steps=sum(r.modes * r.n_steps for r in done),
                          device_steps=self.device_steps, peak_active=peak_active,
                          peak_bytes=peak_bytes)
        return rep

    def fp64_checks(self, report: ServeReport):
        """Re-solve every ``check_every``-th request alone in FP64 (the engine pinned to
        the FP64 oracle at 1e-9) and record the relative L2 of the served result."""
        for r in report.requests:
            if r.rid % self.check_every != 0 or r.result is None:
                continue
            if r.plain:
                r.fp64_rel_l2 = _plain_fp64_rel_l2(r)
                continue
            ref = BatchEngine([r.ctx], precision="fp64", n_steps=r.n_steps)
            ref.run(r.n_steps)
            ref.synchronize()
            b = ref.state(0)
            a = r.result.double()
            r.fp64_rel_l2 = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
            ref.close()
        return [r.fp64_rel_l2 for r in report.requests if not np.isnan(r.fp64_rel_l2)]

    def close(self):
        self.pool.close()


def _plain_fp64_rel_l2(r):
    """Relative L2 of a plain-lane result against the same request re-solved alone with
    the FP64 plain solver (pinned to the FP64 oracle at 1e-9)."""
    import copy

    from ..backends import get_backend
    from ..solvers.spectral import SpectralSolver
    import numpy as _np
    ctx = copy.deepcopy(r.ctx)
    ctx.num.dtype = _np.float64          # the served run's constants, FP64 arithmetic
    if getattr(ctx.num, "transform", "mfma") == "bf16x3":
        ctx.num.transform = "mfma"
    s = SpectralSolver(get_backend("mi355x"), ctx)
    s.initialize()
    dt = ctx.num.dt
    for k in range(r.n_steps):
        s.step(k * dt, dt)
    b = s.state.a.double()
    a = r.result.double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

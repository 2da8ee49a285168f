"""Autotuner for the batched engine's surface transforms.

The batched step is dominated by the surface-transform launches (two per Picard
iteration), or, in the resident variant (rows = −1), by one launch per step in which a
workgroup runs a whole cuboid's Picard loop. Their launch shape has three knobs: rows of the surface per work item
(the MFMA row tiles one workgroup owns: 16 or 32), threads per workgroup (waves per CU:
256, 512 or 1024) and LDS bytes per workgroup (how much of each operand is staged at
once; up to 159 KiB of gfx950's 160 KiB per workgroup, and a smaller budget lets two
workgroups share a CU). The best choice depends on the batch: the number of work items
against the 256 CUs, the widest surface axis, the data type.

``autotune`` builds a scratch engine over the same cuboids, and for every variant the
library offers (``sgh_batch_variants``), every LDS budget and, where the variant has one and
the work list fits one workgroup per CU, the persistent one-launch-per-step kernel as well
as the per-pass launches: rewinds to the initial condition, captures the step graph,
replays ``warmup`` steps, then times ``steps`` graph-replayed steps with HIP events on the
engine's stream (best of ``reps``). Every
variant runs the identical trajectory, so the work compared is the same. The fastest is
returned (and stored in a JSON cache keyed by device, precision and the sorted cuboid
shapes when ``cache_path`` is given).
"""

from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
from dataclasses import asdict, dataclass, field

import torch

from .. import _lib

__all__ = ["LDS_CHOICES", "TuneResult", "autotune", "surface_variants", "signature"]

LDS_CHOICES = (150 * 1024, 159 * 1024, 72 * 1024)


def surface_variants():
    """(rows per item, threads per workgroup) pairs the library was built with."""
    lib = _lib.lib.load()
    n = lib.sgh_batch_variants(None, 0)
    buf = (C.c_int * (2 * n))()
    lib.sgh_batch_variants(C.cast(buf, C.c_void_p), n)
    return [(buf[2 * i], buf[2 * i + 1]) for i in range(n)]


def signature(contexts, precision):
    shapes = sorted(tuple(int(v) for v in c.geom.n) for c in contexts)
    digest = hashlib.sha1(repr((precision, shapes)).encode()).hexdigest()[:16]
    return f"{torch.cuda.get_device_name(0)}|{precision}|{len(shapes)}|{digest}"


@dataclass
class TuneResult:
    rows: int
    threads: int
    lds_bytes: int
    ms_per_step: float
    default_ms_per_step: float
    table: list = field(default_factory=list)
    loop: bool = False

    @property
    def variant(self):
        return (self.rows, self.threads, self.lds_bytes)

    def summary(self):
        return {"rows": self.rows, "threads": self.threads, "lds_bytes": self.lds_bytes,
                "loop": self.loop, "ms_per_step": round(self.ms_per_step, 4),
                "default_ms_per_step": round(self.default_ms_per_step, 4),
                "candidates": len(self.table)}


def _time_steps(eng, steps, reps):
    best = float("inf")
    for _ in range(reps):
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(eng.stream)
        eng.run(steps, use_graph=True)
        t1.record(eng.stream)
        t1.synchronize()
        best = min(best, t0.elapsed_time(t1) / steps)
    return best


def autotune(contexts, precision="mixed", steps=4, warmup=1, reps=2, lds_choices=LDS_CHOICES,
             cache_path=None) -> TuneResult:
    """Time every surface-transform variant on ``contexts`` and return the fastest."""
    from .batch import BatchEngine

    key = signature(contexts, precision)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            return TuneResult(**cache[key])
    per_run = warmup + reps * steps
    eng = BatchEngine(contexts, precision=precision, n_steps=per_run)
    default = eng.variant
    table = []
    try:
        for rows, threads in surface_variants():
            # the resident step kernel (rows < 0) stages nothing by the LDS budget; the budget
            # only shapes its per-pass fallback, so one choice is timed
            for lds in (lds_choices[:1] if rows < 0 else lds_choices):
                try:
                    eng.set_variant(rows, threads, lds)
                except _lib.SGHError:
                    continue   # this LDS budget cannot stage a full-width operand
                # the persistent step kernel is switched off in the library (set_loop(True)
                # is refused), so only the per-pass launches are timed
                for loop in (False,):
                    eng.set_loop(loop)
                    eng.rewind()
                    eng.run(warmup, use_graph=True)
                    eng.synchronize()
                    if loop and not eng.loop_state["active"]:
                        continue   # no persistent kernel for this variant or work list
                    ms = _time_steps(eng, steps, reps)
                    table.append({"rows": rows, "threads": threads, "lds_bytes": lds,
                                  "loop": loop, "ms_per_step": ms})
            eng.set_loop(False)
    finally:
        eng.close()
    best = min(table, key=lambda r: r["ms_per_step"])
    dflt = [r for r in table if (r["rows"], r["threads"], r["lds_bytes"], r["loop"]) ==
            (default["rows"], default["threads"], default["lds_bytes"], False)]
    res = TuneResult(rows=best["rows"], threads=best["threads"], lds_bytes=best["lds_bytes"],
                     ms_per_step=best["ms_per_step"],
                     default_ms_per_step=dflt[0]["ms_per_step"] if dflt else float("nan"),
                     table=table, loop=bool(best["loop"]))
    if cache_path:
        cache[key] = asdict(res)
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return res

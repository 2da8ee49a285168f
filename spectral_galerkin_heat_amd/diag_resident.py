"""Per-phase cycle table of the resident step kernel (a diagnostic build of the library).

On the CPU, once:   python -m spectral_galerkin_heat_amd.diag_resident --build
On the MI355X:      python -m spectral_galerkin_heat_amd.diag_resident [--steps N]

``--build`` compiles the library with ``SGH_RESIDENT_TIMESTAMPS`` into ``_diag/`` next to
this file (the product library is not touched). The run steps the 64-cuboid benchmark batch
(``batch_uneven_cuboids(n=64, seed=0)``) with the resident variant and prints, from the
per-workgroup ``s_memtime`` stamps of the last launch, a markdown table: the phases of the
slowest workgroup (step preparation, warm start, the mean of each Picard stage and of the
team's two field exchanges over the iterations after the first, step end) and the total
cycles per surface shape and team size.
"""

from __future__ import annotations

import argparse
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DIAG_DIR = os.path.join(HERE, "_diag")
SLOTS, BLOCKS = 158, 1024
STAGES = ("inverse 1", "inverse 2 + evaporation", "q exchange", "forward 1",
          "forward 2 + relaxation", "X exchange", "norm sums")
PER = len(STAGES)
END, NX, NY = 143, 144, 145


def build():
    from . import _build
    return _build.build(defines=("SGH_RESIDENT_TIMESTAMPS",), libdir=DIAG_DIR)


def run(steps, team_max=8):
    import numpy as np
    from . import _lib
    _lib.LIB_PATH = os.path.join(DIAG_DIR, "libsgh_mi355x.so")
    from .engine.batch import BatchEngine
    from .workloads import batch_uneven_cuboids, make_context
    cases = batch_uneven_cuboids(n=64, seed=0, n_steps=steps)
    eng = BatchEngine([make_context(c, lp) for c, lp in cases], precision="mixed",
                      variant=(-1, 512, 150 * 1024))
    eng.set_team_max(team_max)
    eng.run(steps)
    eng.synchronize()
    lib = _lib.lib.load()
    fn = lib.sgh_debug_resident_timestamps
    fn.argtypes = [C.c_void_p]
    buf = np.zeros(BLOCKS * SLOTS, dtype=np.uint64)
    if fn(buf.ctypes.data) != 0:
        raise RuntimeError("sgh_debug_resident_timestamps failed")
    met = eng.metrics()
    eng.close()
    ts = buf.reshape(BLOCKS, SLOTS).astype(np.int64)
    live = [b for b in range(BLOCKS) if ts[b, 0] > 0 and ts[b, END] > ts[b, 0]]
    tot = {b: int(ts[b, END] - ts[b, 0]) for b in live}
    worst = max(live, key=lambda b: tot[b])
    t = ts[worst]
    nx, tp = int(t[NX]) & 0xffff, int(t[NX]) >> 16
    ny, mp = int(t[NY]) & 0xffff, int(t[NY]) >> 16
    last = 3 + PER - 1
    n_it = int(sum(1 for k in range(20) if t[last + PER * k] > 0 and t[last + PER * k] >= t[2]))
    lines = [f"Slowest workgroup: surface {nx}x{ny}, member {mp} of a team of {tp}, {n_it} "
             f"Picard iterations recorded (metrics: mean {met[-1, :, 2].mean():.2f} per cuboid "
             f"in the last step), {tot[worst]} cycles in all; {len(live)} workgroups.", "",
             "| phase | cycles | share |", "|---|---:|---:|"]

    def row(name, c):
        lines.append(f"| {name} | {c:.0f} | {100.0 * c / tot[worst]:.1f} % |")

    row("step preparation (AA, laser flux, shifted warm-start flux, q exchange)", t[1] - t[0])
    row("warm start (forward transform, G, Xinv, X exchange)", t[2] - t[1])
    per = {s: [] for s in STAGES}
    prev = None
    for k in range(n_it):
        e = [t[3 + PER * k + j] for j in range(PER)]
        start = t[2] if k == 0 else t[last + PER * (k - 1)]
        if k >= 1:
            for j, s in enumerate(STAGES):
                per[s].append(e[j] - (start if j == 0 else e[j - 1]))
        prev = e[PER - 1]
    iter_total = (prev - t[2]) if prev is not None else 0
    for s in STAGES:
        if per[s]:
            row(f"Picard {s} (mean of iterations 2..{n_it})", float(np.mean(per[s])))
    row(f"all {n_it} Picard iterations", iter_total)
    row("step end (inverse, T max, write-back)", t[END] - (prev if prev is not None else t[2]))
    lines += ["", "Total cycles per surface shape and team size (mean over the workgroups):", "",
              "| nx × ny | team | workgroups | cycles |", "|---|---:|---:|---:|"]
    shapes = {}
    for b in live:
        key = (int(ts[b, NX]) & 0xffff, int(ts[b, NY]) & 0xffff, int(ts[b, NX]) >> 16)
        shapes.setdefault(key, []).append(tot[b])
    for (sx, sy, st), v in sorted(shapes.items(), key=lambda kv: -np.mean(kv[1])):
        lines.append(f"| {sx} × {sy} | {st} | {len(v)} | {np.mean(v):.0f} |")
    # inside two stages of iteration 5 (wave 0's stamps): entry, fragments in registers,
    # tiles done, barrier passed
    lines += ["", "Inside the stages of iteration 6 (wave 0, slowest workgroup): "
              "fragments wait / tiles / barrier wait, cycles"]
    for name, base in (("inverse 1", 146), ("forward 2 + relaxation", 150)):
        e = [int(t[base + j]) for j in range(4)]
        if min(e) > 0:
            lines.append(f"* {name}: {e[1] - e[0]} / {e[2] - e[1]} / {e[3] - e[2]}")
    # inside the step preparation: AA, laser factors, flux pass, laser power, q exchange
    p = [int(t[0])] + [int(t[s]) for s in (154, 155, 156, 157)] + [int(t[1])]
    if min(p) > 0:
        lines += ["", "Inside the step preparation (slowest workgroup): AA sum / laser factors / "
                  "flux pass / laser power / q exchange, cycles",
                  "* " + " / ".join(str(p[k + 1] - p[k]) for k in range(5))]
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--build", action="store_true", help="compile the diagnostic library")
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--team-max", type=int, default=8)
    a = ap.parse_args()
    if a.build:
        print(build())
        return
    print(run(a.steps, a.team_max), flush=True)


if __name__ == "__main__":
    sys.exit(main())

"""Benchmark: mode-steps/s of the batched mixed-precision engine vs the plain FP64 GPU path.

Workload: 64 independent cuboids, seeded (numpy default_rng(0)); per-axis mode counts from
{32,48,64,96,128} (z from {16,24,32}); extents 0.5-4 mm (x, y), 0.25-1 mm (z); P 100-400 W,
radius 35-70 µm, speed 0.2-1.5 m/s, one track at a random y in the middle 60 %; 316L-like
constants, evaporation only; Δt = 5e-6 s; Picard tol 1e-8, cap 20, ω = 0.1.

A step advances all 64 cuboids by one ETD1 step (with its Picard loop). ``value`` is
Σ_i N_i × steps / seconds of the mixed-precision batched engine with its state resident in
HBM, timed between device synchronisations. Printed alongside (in ``config``): the plain
FP64 GPU path (SpectralSolver on the mi355x backend, cuboid by cuboid, one residual read
per Picard iteration), the FP64 batched engine, the relative L2 of the mixed result against
the FP64 batched result (itself pinned to the FP64 oracle at 1e-9 by the parity tests),
a roofline entry for the dominant kernel and the FP64 NumPy oracle timed on host cores.
Before timing, the autotuner picks each engine's surface-transform launch shape. After it,
a continuous-batching request loop serves Poisson arrivals of scan requests (p50/p99
completion latency, aggregate mode-steps/s, FP64 re-solves of every 20th request) and a
sweep of arrival rates ends with a small pool in which admission by footprint queues.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "mode-steps/s vs plain FP64 GPU at ≤1e-4 rel L2; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
N_CU = 256


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=25)
    ap.add_argument("--cpu-cuboids", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-sweep", default="40,160,640,2560",
                    help="comma-separated arrival rates (Hz) for the latency/throughput sweep; "
                         "empty to skip")
    ap.add_argument("--serve-tight-mib", type=int, default=48,
                    help="pool size of the last sweep point, small enough that admission "
                         "by memory footprint has to queue requests")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-plain", type=int, default=10,
                    help="plain-lane requests (latent heat, convection, T-dependent "
                         "properties, box mode) mixed into a 100-request serving point")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--box-steps", type=int, default=40,
                    help="steps of the box-mode single track (80x48x120 modes, fine box "
                         "160x96x112) per precision; 0 skips it")
    ap.add_argument("--box-only", action="store_true",
                    help="run only the box-mode leg and print its JSON (for profiling)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size "
                         "kernel that calibrates HBM byte counters under a profiler)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default surface-transform variant instead of autotuning")
    ap.add_argument("--variant", default="",
                    help="ROWS,THREADS,LDS_BYTES: run the mixed engine with this variant instead "
                         "of autotuning it (rows -1: the resident step kernel; for profiling)")
    return ap.parse_args()


def surface_flops_per_launch(ctxs):
    """Algorithmic flops of one batched 2-D transform launch: Σ 2·S·(nx + ny)."""
    tot = 0
    for c in ctxs:
        nx, ny = int(c.geom.n.x), int(c.geom.n.y)
        tot += 2 * nx * ny * (nx + ny)
    return tot


def resident_flops_per_launch(ctxs, iters):
    """bf16 MFMA flops of one resident step launch: per cuboid 2 + 2·n_it 2-D transforms
    (warm start forward, inverse + forward per Picard iteration, step-end inverse), each
    2·S·(nx + ny) FP32-equivalent flops done as three bf16 products (hi·hi + hi·lo + lo·hi)."""
    tot = 0.0
    for c, it in zip(ctxs, iters):
        nx, ny = int(c.geom.n.x), int(c.geom.n.y)
        tot += 3 * 2 * nx * ny * (nx + ny) * (2 + 2 * float(it))
    return tot


def surface_bytes_per_launch(ctxs, word):
    """Algorithmic HBM bytes of one Picard surface launch, the mean of the inverse
    launch (reads Xinv, q_las; writes q_evap, q_net: 4 S-words) and the forward launch
    (reads q_net, G, Wtt, U, ZE, Ztt; writes G, Xinv: 8 S-words): 6·S words."""
    return sum(6 * word * int(c.geom.n.x) * int(c.geom.n.y) for c in ctxs)


def volume_bytes_per_launch(ctxs, word):
    """Algorithmic HBM bytes of one volume pass: read + write the state, plus the
    surface reads/writes (G in, ZE and U out, q_evap copied)."""
    tot = 0
    for c in ctxs:
        nx, ny, nz = int(c.geom.n.x), int(c.geom.n.y), int(c.geom.n.z)
        tot += 2 * word * nx * ny * nz + 5 * word * nx * ny
    return tot


def box_mode_leg(steps, with_cpu):
    """Plain device path with a laser-following fine box (the reference's example-02 single
    track): FP64 and FP32 rates, FP32 vs FP64 after the same steps, the FP64 state vs the
    oracle after one step, and the oracle's CPU time for that step."""
    import numpy as np
    import torch
    from spectral_galerkin_heat_amd.backends import get_backend
    from spectral_galerkin_heat_amd.physics.meltpool import melt_pool_extent
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as K
    from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
    from spectral_galerkin_heat_amd.workloads import make_context, single_track_box
    res, states = {}, {}
    for dtype in ("float64", "float32"):
        ctx = make_context(*single_track_box(n_steps=steps + 2, dtype=dtype))
        s = SpectralSolver(get_backend("mi355x"), ctx)
        s.initialize()
        dt = ctx.num.dt
        for n in range(2):   # untimed: first-call set-up
            s.step(n * dt, dt)
        its = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for n in range(2, 2 + steps):
            its.append(s.step(n * dt, dt)[1]["n_evap_iter"])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        N = int(np.prod(tuple(ctx.geom.n)))
        T = K.reconstruct_volume(s.state.a, s.state)
        pool = melt_pool_extent(T, tuple(ctx.geom.d), T_melt=float(ctx.mat.T_liquidus))
        states[dtype] = s.state.a.double().clone()
        res[dtype] = {"ms_per_step": round(el / steps * 1e3, 3),
                      "mode_steps_per_s": round(N * steps / el, 1),
                      "picard_iterations_mean": round(float(np.mean(its)), 2),
                      "melt_pool_um": [round(v * 1e6, 2) for v in pool]}
        log(f"box mode {dtype}: {res[dtype]}")
    ref = states["float64"]
    out = {"workload": "box-mode single track, 80x48x120 modes, fine box 160x96x112 "
                       f"(refinement 2, sliding in x), latent heat + bottom convection, steps 2..{steps + 1} timed",
           "fp64": res["float64"], "fp32": res["float32"],
           "rel_l2_fp32_vs_fp64": float(f"{float(torch.linalg.vector_norm(states['float32'] - ref) / torch.linalg.vector_norm(ref)):.3e}")}
    if with_cpu:
        from oracle.sgh_oracle import OracleSolver
        ctx = make_context(*single_track_box(n_steps=1, dtype="float64"))
        s = SpectralSolver(get_backend("mi355x"), ctx)
        s.initialize()
        s.step(0.0, ctx.num.dt)
        import oracle.sgh_oracle as orc
        from threadpoolctl import threadpool_limits
        N = int(np.prod(tuple(ctx.geom.n)))
        a = s.state.a.double().cpu().numpy()
        t_cpu = {}
        for label, workers, blas in (("all_cores", -1, None), ("one_core", 1, 1)):
            r = OracleSolver(make_context(*single_track_box(n_steps=1))).initialize()
            saved = orc.FFT_WORKERS
            orc.FFT_WORKERS = workers
            try:
                with threadpool_limits(limits=blas):
                    t0 = time.perf_counter()
                    r.step(0.0, ctx.num.dt)
                    t_cpu[label] = time.perf_counter() - t0
            finally:
                orc.FFT_WORKERS = saved
        best = min(t_cpu.values())
        out["fp64_vs_oracle_rel_l2_step1"] = float(f"{np.linalg.norm(a - r.a) / np.linalg.norm(r.a):.3e}")
        out["cpu_oracle_mode_steps_per_s"] = round(N / best, 1)
        out["cpu_oracle_seconds"] = {k: round(v, 3) for k, v in t_cpu.items()}
        out["fp64_speedup_vs_cpu_oracle"] = round(res["float64"]["mode_steps_per_s"] / (N / best), 1)
    return out


def measured_stream_gbps(mib=1024, reps=5):
    """Read + write bandwidth of a plain device copy of `mib` MiB (best of `reps`)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        dst.copy_(src)
        t1.record()
        t1.synchronize()
        best = min(best, t0.elapsed_time(t1) * 1e-3)
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * mib * (1 << 20) / best / 1e9


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU engine: --gpus must be 1. Independent "
                          "simulations shard by request (one pool per GPU, no collectives); "
                          "multi-GPU runs are not part of this package"}))
        sys.exit(2)
    import logging

    import numpy as np
    import torch

    logging.getLogger("spectral_galerkin_heat_amd").setLevel(logging.ERROR)
    from spectral_galerkin_heat_amd.engine.batch import BatchEngine
    from spectral_galerkin_heat_amd.workloads import batch_uneven_cuboids, make_context

    assert torch.cuda.is_available(), "bench.py needs the MI355X"
    if args.box_only:
        print(json.dumps({"box_mode": box_mode_leg(max(1, args.box_steps), not args.no_cpu)}),
              flush=True)
        return
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del src, dst
    W, K = args.warmup, args.steps
    instr = min(K, 40)
    total_steps = W + K + instr
    cases = batch_uneven_cuboids(n=64, seed=0, n_steps=total_steps)
    ctxs = [make_context(c, lp) for c, lp in cases]
    N_total = sum(int(np.prod(tuple(c.geom.n))) for c in ctxs)
    log(f"workload: 64 cuboids, sum N = {N_total}, warmup {W}, timed {K}")

    tuning = {}
    if not args.no_tune:
        from spectral_galerkin_heat_amd.engine.autotune import autotune
        for prec in (("fp64",) if args.variant else ("mixed", "fp64")):
            tuning[prec] = autotune(ctxs, prec)
            log(f"autotune {prec}: {tuning[prec].summary()}")
            for row in sorted(tuning[prec].table, key=lambda r: r["ms_per_step"]):
                log(f"    rows {row['rows']:2d} threads {row['threads']:4d} lds {row['lds_bytes'] >> 10:3d} KiB"
                    f"{' persistent' if row.get('loop') else ''}: {row['ms_per_step']:.4f} ms/step")

    forced = tuple(int(v) for v in args.variant.split(",")) if args.variant else None

    def timed_engine(precision):
        tv = tuning.get(precision)
        variant = tv.summary() if tv else None
        if forced is not None and precision == "mixed":
            variant = forced
        eng = BatchEngine(ctxs, precision=precision, variant=variant)
        eng.run(W)
        eng.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run(K)
        eng.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return eng, dt

    eng_mx, t_mx = timed_engine("mixed")
    value = N_total * K / t_mx
    log(f"mixed batched: {t_mx * 1e3:.1f} ms for {K} steps -> {value:.3e} mode-steps/s")
    met = eng_mx.metrics()
    n_it_mean = float(met[W:, :, 2].mean())

    eng_64, t_64 = timed_engine("fp64")
    fp64_rate = N_total * K / t_64
    log(f"fp64 batched: {t_64 * 1e3:.1f} ms -> {fp64_rate:.3e} mode-steps/s")
    from types import SimpleNamespace

    from spectral_galerkin_heat_amd.io_utils import compute_L2_structured
    from spectral_galerkin_heat_amd.physics.meltpool import cell_temperature, melt_pool_extent
    from spectral_galerkin_heat_amd.physics.spectral_helpers import reconstruct_temperature_DCT
    num, den, num_x, den_x = 0.0, 0.0, 0.0, 0.0
    rel_trap = []   # the reference's metric: trapezoid-weighted rel-L2 on the node fields
    rel_max = 0.0
    pool_dev = 0.0   # largest |Δwidth|/dy or |Δdepth|/dz over the cuboids
    melting = 0
    for i in range(64):
        a = eng_mx.state(i).double()
        b = eng_64.state(i)
        d2 = float(torch.sum((a - b) ** 2))
        n2 = float(torch.sum(b ** 2))
        num += d2
        den += n2
        rel_max = max(rel_max, (d2 / n2) ** 0.5)
        # excursion T − T0: the mean mode without T0·√V
        g = ctxs[i].geom
        m0 = float(ctxs[i].mat.T0) * float(np.sqrt(float(g.size.x) * float(g.size.y) * float(g.size.z)))
        num_x += d2
        den_x += n2 - float(b[0, 0, 0]) ** 2 + (float(b[0, 0, 0]) - m0) ** 2
        d = tuple(float(v) for v in g.d)
        pa = melt_pool_extent(cell_temperature(a, g), d)
        pb = melt_pool_extent(cell_temperature(b, g), d)
        melting += pb[1] > 0
        pool_dev = max(pool_dev, abs(pa[1] - pb[1]) / d[1], abs(pa[2] - pb[2]) / d[2])
        shim = SimpleNamespace(grid=SimpleNamespace(dx=d[0], dy=d[1], dz=d[2]))
        Tb = reconstruct_temperature_DCT(b, shim).permute(2, 1, 0)
        Ta = reconstruct_temperature_DCT(a, shim).permute(2, 1, 0)
        nz_, ny_, nx_ = b.shape
        coords = [np.linspace(0.0, n * h, n + 1) for n, h in ((nx_, d[0]), (ny_, d[1]), (nz_, d[2]))]
        rel_trap.append(compute_L2_structured(Tb, Ta, *coords)["L2_rel"])
    rel_batch = (num / den) ** 0.5
    rel_exc = (num_x / den_x) ** 0.5
    log(f"mixed vs fp64: batch rel-L2 {rel_batch:.2e} (excursion {rel_exc:.2e}), worst cuboid "
        f"{rel_max:.2e}; melt pools in {melting} cuboids, max width/depth deviation "
        f"{pool_dev:.2f} cells; node fields, trapezoid weights: worst {max(rel_trap):.2e}")
    eng_64.close()

    # live per-kernel timing (HIP events on the engine stream) over an instrumented segment
    # of the same workload (steps W+K .. W+K+instr, all cuboids still live)
    eng_mx.set_timing(True)
    eng_mx.run(instr, use_graph=False)
    eng_mx.synchronize()
    tm = eng_mx.timing()
    eng_mx.set_timing(False)
    surf_ms = tm["surface_ms"] / max(tm["surface_launches"], 1)
    vol_ms = tm["volume_ms"] / max(tm["volume_launches"], 1)
    step_ms = t_mx / K * 1e3
    share_surf = tm["surface_ms"] / instr / step_ms
    share_vol = tm["volume_ms"] / instr / step_ms
    log(f"kernel avg: surface {surf_ms * 1e3:.1f} us ({share_surf:.0%} of step), "
        f"volume {vol_ms * 1e3:.1f} us ({share_vol:.0%})")
    flops = surface_flops_per_launch(ctxs)
    sbytes = surface_bytes_per_launch(ctxs, 4)
    vbytes = volume_bytes_per_launch(ctxs, 4)
    resident = tm.get("surface_launches", 0) == instr and eng_mx.variant["rows"] < 0
    if resident and share_surf >= share_vol:
        # one launch per step, a team of workgroups (one CU each) per cuboid: the bound is
        # the matrix cores of the CUs it occupies; iterations per cuboid from the same
        # segment's metrics
        iters = eng_mx.metrics()[W + K:W + K + instr, :, 2].mean(axis=0)
        rflops = resident_flops_per_launch(ctxs, iters)
        tfs = rflops / (surf_ms * 1e-3) / 1e12
        teams = eng_mx.team_sizes
        active = min(max(sum(teams), len(ctxs)), N_CU) / N_CU
        roofline = {"kernel": "k_b_resident (a team of 1-8 workgroups per cuboid: laser flux, "
                              "warm start, every Picard iteration and the step end in one "
                              "launch; split-bf16 MFMA transforms)",
                    "bound": "mfma", "achieved": round(tfs, 3), "peak": BF16_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(tfs / BF16_PEAK_TFLOPS, 5),
                    "traffic": None, "avg_launch_us": round(surf_ms * 1e3, 2),
                    "bf16_flops_per_launch": rflops,
                    "picard_iterations_per_cuboid_mean": round(float(iters.mean()), 2),
                    "active_cu_fraction": round(active, 4),
                    "team_sizes": sorted((t for t in teams if t > 0), reverse=True),
                    "frac_of_active_cus": round(tfs / (BF16_PEAK_TFLOPS * active), 5)}
    elif share_surf >= share_vol:
        # the binding roofline of the surface launches is the larger of the two lower
        # bounds on their time: algorithmic bytes at HBM peak, algorithmic flops at MFMA peak
        t_hbm = sbytes / (HBM_PEAK_GBS * 1e9)
        t_mfma = flops / (FP32_PEAK_TFLOPS * 1e12)
        gbs = sbytes / (surf_ms * 1e-3) / 1e9
        tfs = flops / (surf_ms * 1e-3) / 1e12
        kname = "k_b_surface (batched 2-D cosine transform + Picard epilogue)"
        if t_hbm >= t_mfma:
            roofline = {"kernel": kname, "bound": "hbm", "achieved": round(gbs, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5)}
        else:
            roofline = {"kernel": kname, "bound": "mfma", "achieved": round(tfs, 3),
                        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tfs / FP32_PEAK_TFLOPS, 5)}
        roofline.update({"traffic": None, "avg_launch_us": round(surf_ms * 1e3, 2),
                         "algorithmic_bytes_per_launch": sbytes,
                         "algorithmic_flops_per_launch": flops,
                         "mfma_frac": round(tfs / FP32_PEAK_TFLOPS, 5),
                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 5)})
    else:
        achieved = vbytes / (vol_ms * 1e-3) / 1e9
        roofline = {"kernel": "k_b_volume (ETD1 state pass)", "bound": "hbm",
                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_launch_us": round(vol_ms * 1e3, 2),
                    "algorithmic_bytes_per_launch": vbytes}
    vol_gbs = vbytes / (vol_ms * 1e-3) / 1e9
    # measured-stream denominator: a 1 GiB device copy (read + write) timed with events
    stream_gbs = measured_stream_gbps()
    log(f"measured stream copy: {stream_gbs:.0f} GB/s (read + write)")
    if roofline["unit"] == "GB/s":
        roofline["measured_stream_GBps"] = round(stream_gbs, 1)
        roofline["frac_vs_measured_stream"] = round(roofline["achieved"] / stream_gbs, 5)
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        roofline["traffic"] = tr.get(roofline["kernel"].split()[0])
    eng_mx.close()

    # plain FP64 GPU path: SpectralSolver per cuboid
    plain_rate = None
    if not args.no_plain:
        from spectral_galerkin_heat_amd.backends import get_backend
        from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
        backend = get_backend("mi355x")
        solvers = []
        for c, lp in cases:
            s = SpectralSolver(backend, make_context(c, lp))
            s.initialize()
            solvers.append(s)
        for s in solvers:
            s.step(0.0, s.context.num.dt)
        torch.cuda.synchronize()
        P = args.plain_steps
        t0 = time.perf_counter()
        for s in solvers:
            dt = s.context.num.dt
            for n in range(1, 1 + P):
                s.step(n * dt, dt)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        plain_rate = N_total * P / t_plain
        log(f"plain fp64 path: {t_plain:.2f} s for {P} steps x 64 cuboids -> {plain_rate:.3e}")
        del solvers

    box = None
    if args.box_steps > 0:
        box = box_mode_leg(args.box_steps, not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        # the oracle on the host cores, twice over the same sample: scipy.fft on every core
        # (the reference's workers=-1) and on one core with one BLAS thread; the faster is
        # the baseline (threading does not pay at these transform sizes)
        import oracle.sgh_oracle as orc
        from oracle.sgh_oracle import OracleSolver
        from threadpoolctl import threadpool_limits
        cores = len(os.sched_getaffinity(0))
        nc, ns = args.cpu_cuboids, args.cpu_steps

        def time_oracle(workers, blas_threads):
            refs = [OracleSolver(make_context(c, lp)).initialize() for c, lp in cases[:nc]]
            saved = orc.FFT_WORKERS
            orc.FFT_WORKERS = workers
            try:
                with threadpool_limits(limits=blas_threads):
                    t0 = time.perf_counter()
                    for r in refs:
                        dt = r.ctx.num.dt
                        for n in range(ns):
                            r.step(n * dt, dt)
                    t = time.perf_counter() - t0
            finally:
                orc.FFT_WORKERS = saved
            return sum(int(r.a.size) for r in refs) * ns / t, t

        rate_all, t_all = time_oracle(-1, None)
        log(f"cpu oracle: {t_all:.1f} s -> {rate_all:.3e} mode-steps/s on {cores} cores")
        rate_one, t_one = time_oracle(1, 1)
        log(f"cpu oracle, one core: {t_one:.1f} s -> {rate_one:.3e} mode-steps/s")
        best_one = rate_one >= rate_all
        sample = (f"FP64 NumPy/SciPy oracle (restated reference step, no Numba), first {nc} of "
                  f"the 64 cuboids x {ns} steps; {cores} cores (scipy.fft workers=-1): "
                  f"{rate_all:.4g} mode-steps/s in {t_all:.1f} s; one core (workers=1, one BLAS "
                  f"thread): {rate_one:.4g} in {t_one:.1f} s; value = the faster")
        cpu = {"value": round(max(rate_all, rate_one), 1), "unit": "mode-steps/s",
               "cores": 1 if best_one else cores, "kind": "port", "sample": sample,
               "cpu_model": cpu_model(),
               "all_cores": {"cores": cores, "value": round(rate_all, 1), "seconds": round(t_all, 2)},
               "one_core": {"cores": 1, "value": round(rate_one, 1), "seconds": round(t_one, 2)}}

    serving = None
    sweep = None
    if not args.no_serve:
        from spectral_galerkin_heat_amd.engine.serving import Request, ScanServer, poisson_arrivals
        from spectral_galerkin_heat_amd.workloads import serving_request
        rng = np.random.default_rng(1)
        shapes = [serving_request(rng) for _ in range(args.serve_requests)]
        big = sum(1 for cfg, _ in shapes if np.prod(cfg["domain"]["mesh"]) >= 256 * 128 * 64)

        def serve(rate, budget, checks, shapes=shapes):
            arr = poisson_arrivals(rate, len(shapes), np.random.default_rng(2))
            reqs = []
            for i, (cfg, lp) in enumerate(shapes):
                ctx = make_context(cfg, lp)
                reqs.append(Request(rid=i, ctx=ctx, n_steps=int(ctx.num.n_steps),
                                    arrival=float(arr[i])))
            srv = ScanServer(precision="mixed", max_cuboids=256, hbm_budget_bytes=budget,
                             max_steps=2048, steps_per_tick=8, check_every=20)
            rep = srv.serve(reqs)
            fp64 = srv.fp64_checks(rep) if checks else []
            srv.close()
            sm = rep.summary()
            out = {"arrival_rate_hz": rate, "requests": sm["requests"],
                   "hbm_budget_mib": budget >> 20,
                   "p50_latency_s": round(sm["p50_latency_s"], 4),
                   "p99_latency_s": round(sm["p99_latency_s"], 4),
                   "aggregate_mode_steps_per_s": round(sm["aggregate_mode_steps_per_s"], 1),
                   "queue_wait_p50_s": round(sm["queue_wait_p50_s"], 4),
                   "queue_wait_max_s": round(sm["queue_wait_max_s"], 4),
                   "peak_active": sm["peak_active"],
                   "peak_pool_mib": round(sm["peak_pool_bytes"] / 2**20, 1),
                   "wall_s": round(sm["wall_s"], 3)}
            if sm["plain_requests"]:
                pl = [r.latency for r in rep.requests if r.plain]
                out["plain_requests"] = sm["plain_requests"]
                out["plain_p50_latency_s"] = round(float(np.percentile(pl, 50)), 4)
                out["plain_max_latency_s"] = round(float(np.max(pl)), 4)
            if checks:
                out["fp64_recheck_rel_l2"] = [float(f"{c:.3e}") for c in fp64]
            return out

        serving = serve(args.serve_rate, 8 << 30, True)
        serving["large_256x128x64_requests"] = big
        log(f"serving: {serving}")
        if args.serve_sweep:
            sweep = []
            rates = [float(r) for r in args.serve_sweep.split(",")]
            for rate in rates:
                sweep.append(serve(rate, 8 << 30, False))
                log(f"sweep: {sweep[-1]}")
            # the same load with a small pool: admission by footprint saturates and queues
            sweep.append(serve(rates[-1], args.serve_tight_mib << 20, False))
            log(f"sweep (tight pool): {sweep[-1]}")
        if args.serve_plain > 0:
            # production physics beside the batched requests: every tenth of the first 100
            # requests replaced by a plain-lane case (T-dependent track with latent heat and
            # convection, FP32; every fifth a box-mode track, FP64)
            from spectral_galerkin_heat_amd.workloads import single_track_box, single_track_tdep
            prng = np.random.default_rng(3)
            mixed = list(shapes[:100])
            for j in range(min(args.serve_plain, 10)):
                n = int(prng.integers(40, 101))
                if j % 5 == 4:
                    case = single_track_box(mesh=(40, 24, 60), n_steps=n,
                                            box_size=(0.3e-3, 0.3e-3, 0.05e-3))
                else:
                    case = single_track_tdep(mesh=(64, 32, 16), n_steps=n, dtype="float32")
                mixed[10 * j] = case
            serving_mixed = serve(args.serve_rate * 2, 8 << 30, True, shapes=mixed)
            log(f"serving (mixed physics): {serving_mixed}")
            serving["mixed_physics"] = serving_mixed

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "mode-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 (state, fields) + split-bf16 MFMA transforms with f32 accumulation + f64 "
                  "(E/KK tables, Picard norms, reductions)" if roofline["kernel"].startswith("k_b_resident")
                  else "f32 (state, transforms, forcing) + f64 (E/KK tables, Picard norms, reductions)"),
        "data": "synthetic",
        "config": {
            "workload": "64 uneven cuboids in one batch (seed 0), constant-property single "
                        "tracks with evaporation, dt 5e-6 s, Picard tol 1e-8 cap 20, omega 0.1",
            "precision": "mixed",
            "sum_modes": N_total,
            "picard_iterations_mean": round(n_it_mean, 2),
            "plain_fp64_gpu_mode_steps_per_s": None if plain_rate is None else round(plain_rate, 1),
            "speedup_vs_plain_fp64_gpu": None if plain_rate is None else round(value / plain_rate, 2),
            "fp64_batched_mode_steps_per_s": round(fp64_rate, 1),
            "rel_l2_mixed_vs_fp64_batch": float(f"{rel_batch:.3e}"),
            "rel_l2_mixed_vs_fp64_worst_cuboid": float(f"{rel_max:.3e}"),
            "rel_l2_excursion_mixed_vs_fp64_batch": float(f"{rel_exc:.3e}"),
            "rel_l2_mixed_vs_fp64_node_trapezoid_worst_cuboid": float(f"{max(rel_trap):.3e}"),
            "rel_l2_mixed_vs_fp64_node_trapezoid_mean": float(f"{float(np.mean(rel_trap)):.3e}"),
            "melt_pool_cuboids": melting,
            "melt_pool_max_width_depth_deviation_cells": round(pool_dev, 3),
            "volume_pass_GBps": round(vol_gbs, 1),
            "kernel_share_of_step": {"surface_iterations": round(share_surf, 3),
                                     "volume_pass": round(share_vol, 3)},
            "parallelism": "single GPU",
            "autotune": {p: t.summary() for p, t in tuning.items()},
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serving": serving,
        "serving_sweep": sweep,
        "box_mode": box,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

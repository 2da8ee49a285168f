"""The surface-collapsed Picard iteration equals the reference's full-volume iteration.

The batched engine (csrc/sgh_batch.hip) never forms the volume iterate: it carries a
surface field G with a = aE + KK∘(w⊗G) and evaluates the surface temperature, the
residual norm and the iterate norm through the tables ZE, U, AA, Ztt, Wtt. This CPU test
writes that algorithm in NumPy (FP64) and compares every step with the oracle's
full-volume restatement of the reference step: same state to 1e-12 and the same Picard
counts. A GPU mismatch is then a kernel bug, not an algebra error.
"""

import math

import numpy as np

from oracle import sgh_oracle as o
from spectral_galerkin_heat_amd.workloads import batch_uneven_cuboids, make_context, single_track_constant


def collapsed_run(ctx, n_steps):
    g = o.OracleGrid(tuple(ctx.geom.size), tuple(ctx.geom.n))
    mat, num, lp = ctx.mat, ctx.num, ctx.laser
    K, KK = o.precompute_K_KK(float(mat.k), float(mat.rho), float(mat.Cp), float(num.dt),
                              (g.nz, g.ny, g.nx), (g.Lz, g.Ly, g.Lx))
    w = g.Cp_top[:, 0, 0]
    Ztt = np.einsum("p,pij->ij", w * w, KK)
    Wtt = np.einsum("p,pij->ij", w * w, KK * KK)
    Bx = g.C[0][:, None] * np.cos(np.pi * np.arange(g.nx)[:, None] * g.x[None, :] / g.Lx)
    By = g.C[1][:, None] * np.cos(np.pi * np.arange(g.ny)[:, None] * g.y[None, :] / g.Ly)
    a = np.zeros((g.nz, g.ny, g.nx))
    a[0, 0, 0] = float(mat.T0) * math.sqrt(g.Lx * g.Ly * g.Lz)
    q_old = np.zeros((g.ny, g.nx))
    N = float(a.size)
    omega = float(num.picard_omega if num.picard_omega is not None else 0.1)
    tol = float(num.picard_tol)
    cap = int(num.max_picard_iter)
    its = []
    for n in range(n_steps):
        ls = ctx.laser_path.get_state(n * num.dt, num.dt)
        I0 = float(lp.absorptivity) * float(ls.power) / (math.pi / 2 * float(lp.r_x) * float(lp.r_y)) if ls.is_on else 0.0
        q_las = o.super_gaussian_flux(g.x, g.y, float(ls.x), float(ls.y), float(lp.r_x), float(lp.r_y), 2.0, I0)
        aE = a * K
        ZE = np.einsum("p,pij->ij", w, aE)
        U = np.einsum("p,pij->ij", w, aE * KK)
        AA = float(np.sum(aE * aE))
        q_w = q_las - o.shift_flux(q_old, (float(ls.v[0]) * num.dt, float(ls.v[1]) * num.dt), g.dx, g.dy)
        G = g.dx * g.dy * By @ q_w @ Bx.T
        q_evap = np.zeros_like(q_old)
        k_done = 0
        for k in range(cap):
            T_s = By.T @ (ZE + G * Ztt) @ Bx
            q_evap = o.evaporation_flux(T_s, float(mat.Pa), float(mat.T_boil), float(mat.DeltaH_LV),
                                        float(mat.R_v), float(mat.T_liquidus))
            S_raw = g.dx * g.dy * By @ (q_las - q_evap) @ Bx.T
            D = S_raw - G
            rr = float(np.sum(D * D * Wtt))
            aa = AA + 2.0 * float(np.sum(G * U)) + float(np.sum(G * G * Wtt))
            G = G + omega * D
            k_done = k + 1
            rel = math.sqrt(rr / N) / max(math.sqrt(aa / N), 1e-9)
            if rel < tol:
                break
        its.append(k_done)
        a = aE + KK * w[:, None, None] * G[None]
        q_old = q_evap
    return a, its


def _oracle(ctx, n_steps):
    s = o.OracleSolver(ctx).initialize()
    its = []
    for n in range(n_steps):
        _, m = s.step(n * ctx.num.dt, ctx.num.dt)
        its.append(m["n_evap_iter"])
    return s.a, its


def test_collapsed_iteration_equals_volume_iteration_single_track():
    cfg, lp = single_track_constant(n_steps=25)
    a_c, its_c = collapsed_run(make_context(cfg, lp), 25)
    a_o, its_o = _oracle(make_context(cfg, lp), 25)
    assert its_c == its_o
    assert o.modal_rel_l2(a_c, a_o) < 1e-12


def test_collapsed_iteration_with_early_convergence():
    # A loose tolerance makes steps converge before the cap, exercising the break path.
    cases = batch_uneven_cuboids(n=2, seed=3, n_steps=12)
    for cfg, lp in cases:
        cfg["simulation"]["picard_tol"] = 3e-6
        a_c, its_c = collapsed_run(make_context(cfg, lp), 12)
        a_o, its_o = _oracle(make_context(cfg, lp), 12)
        assert its_c == its_o and min(its_o) < 20
        assert o.modal_rel_l2(a_c, a_o) < 1e-12

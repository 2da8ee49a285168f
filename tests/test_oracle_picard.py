"""Pin the oracle's relaxed Picard loop (SpectralSolver.step, sgh/solvers/spectral.py:356-485).

Every headline configuration runs at or near the Picard cap, so a capped step's result
depends on the relaxation, the norm, the convergence decision and the warm start, not
only on the fixed point. These checks fix each of them against a computation the oracle
does not share:

* the converged iterate is a fixed point of the forcing map F, evaluated with brute-force
  cosine sums (not ``dct2``) and the evaporation law written from p_sat and the kinetic
  flux, and the fixed point is the same for two relaxation factors;
* with the cap at 1 and 2 the oracle's iterates, residual norm and stored evaporation
  flux equal a step-by-step computation of ``a_new = a_old + ω·(F(a_old) − a_old)`` and
  ``rel = ‖r‖ / max(‖a_old‖, 1e-9·√N)``, from a warm start whose shifted flux comes from
  ``scipy.ndimage.shift`` by the laser's +v·Δt (the mirrored shift fails the same check);
* the reference's own nonlinear invariants: forced cap
  (tests/integration/test_latent_grid_mode.py:98-110), a monotone residual in a converging
  T-dependent box-mode run (tests/integration/test_property_correction.py:132-170) and the
  physical bounds of a full nonlinear run (tests/integration/test_validation.py:152-189).
"""

import copy
import math

import numpy as np
import pytest
import scipy.ndimage

from oracle import sgh_oracle as o
from spectral_galerkin_heat_amd.core.laser import ConstantVelocityLaser
from spectral_galerkin_heat_amd.workloads import MATERIAL_316L, make_context

R_GAS = 8.314462618

# --- independent forcing map ----------------------------------------------------------


def _cos_basis(n, L):
    """phi[k, i] = C_k cos(kπ x_i / L) at the cell centres x_i = (i + ½)L/n."""
    x = (np.arange(n) + 0.5) * L / n
    k = np.arange(n)[:, None]
    C = np.where(k == 0, math.sqrt(1.0 / L), math.sqrt(2.0 / L))
    return C * np.cos(np.pi * k * x[None, :] / L)


class BruteForce:
    """F(a) = aE + KK·(φ_z(Lz) ⊗ S(q_las − q_evap(T_s(a)))) from explicit sums."""

    def __init__(self, ctx):
        g, m = ctx.geom, ctx.mat
        self.nx, self.ny, self.nz = (int(v) for v in g.n)
        self.Lx, self.Ly, self.Lz = (float(v) for v in g.size)
        self.dx, self.dy = self.Lx / self.nx, self.Ly / self.ny
        self.px, self.py = _cos_basis(self.nx, self.Lx), _cos_basis(self.ny, self.Ly)
        kz = np.arange(self.nz)
        Cz = np.where(kz == 0, math.sqrt(1.0 / self.Lz), math.sqrt(2.0 / self.Lz))
        self.wz = Cz * np.cos(np.pi * kz)                      # φ_p(z = Lz)
        alpha = float(m.k) / (float(m.rho) * float(m.Cp))
        lam = sum(np.meshgrid(*[(np.pi * np.arange(n) / L) ** 2 for n, L in
                                ((self.nz, self.Lz), (self.ny, self.Ly), (self.nx, self.Lx))],
                              indexing="ij"))
        dt = float(ctx.num.dt)
        self.E = np.exp(-alpha * lam * dt)
        with np.errstate(divide="ignore", invalid="ignore"):
            # ∫₀^Δt e^{−αλ(Δt−s)} ds / ρc, and Δt/ρc for the mean mode
            self.KK = np.where(lam > 0, -np.expm1(-alpha * lam * dt) / (alpha * lam), dt)
        self.KK /= float(m.rho) * float(m.Cp)
        # evaporation from p_sat(T) and the kinetic flux √(M/2πRT), M = R/R_v
        self.M = R_GAS / float(m.R_v)
        self.Lv, self.Pa, self.Tb = float(m.DeltaH_LV), float(m.Pa), float(m.T_boil)
        self.T_liq = float(m.T_liquidus)
        self.ctx = ctx

    def surface_T(self, a):
        return np.einsum("pnm,p,nj,mi->ji", a, self.wz, self.py, self.px)

    def project(self, q):
        return np.einsum("ji,nj,mi->nm", q, self.py, self.px) * self.dx * self.dy

    def q_evap(self, T):
        p_sat = self.Pa * np.exp(self.Lv * self.M / R_GAS * (1.0 / self.Tb - 1.0 / T))
        q = 0.82 * self.Lv * p_sat * np.sqrt(self.M / (2.0 * np.pi * R_GAS * T))
        return np.where(T >= self.T_liq, q, 0.0)

    def q_laser(self, t):
        lp, path = self.ctx.laser, self.ctx.laser_path
        st = path.get_state(t, float(self.ctx.num.dt))
        x = (np.arange(self.nx) + 0.5) * self.dx
        y = (np.arange(self.ny) + 0.5) * self.dy
        r2 = ((x[None, :] - st.x) ** 2 + (y[:, None] - st.y) ** 2) / float(lp.r_x) ** 2
        I0 = 2.0 * float(lp.absorptivity) * float(st.power) / (np.pi * float(lp.r_x) ** 2)
        return I0 * np.exp(-2.0 * r2)

    def apply(self, aE, q_surface):
        return aE + self.KK * self.wz[:, None, None] * self.project(q_surface)[None]

    def F(self, aE, a, t):
        return self.apply(aE, self.q_laser(t) - self.q_evap(self.surface_T(a)))


def _evap_case(omega=0.1, tol=1e-13, cap=2000, power=100.0):
    """16×12×8 modes, 0.4×0.3×0.1 mm, 100 W moving at 0.5 m/s, Δt = 1e-5 s: after one step
    the surface is above 3500 K and evaporation takes a large share of the laser flux."""
    mat = dict(MATERIAL_316L)
    cfg = {"simulation": {"dt": 1e-5, "duration": 1e-4, "dtype": "float64",
                          "picard_omega": omega, "picard_tol": tol, "max_picard_iter": cap},
           "domain": {"size": [0.4e-3, 0.3e-3, 0.1e-3], "mesh": [16, 12, 8]}, "material": mat,
           "laser": {"radius": 40e-6, "absorptivity": 0.35, "power_nominal": power}, "io": {}}
    return make_context(cfg, ConstantVelocityLaser(0.1e-3, 0.15e-3, 0.5, 0.0, power))


def _after_first_step():
    s = o.OracleSolver(_evap_case()).initialize()
    s.step(0.0, 1e-5)
    return s


def test_brute_force_map_agrees_with_the_oracle_transforms():
    # the brute-force sums and the oracle's DCT forms are the same operators
    s = _after_first_step()
    bf = BruteForce(s.ctx)
    np.testing.assert_allclose(bf.surface_T(s.a), s.surface_T(), rtol=1e-12)
    q = np.random.default_rng(1).standard_normal((12, 16))
    np.testing.assert_allclose(bf.project(q), s.grid.dct_scale * o.dct2(q), atol=1e-12 *
                               np.abs(bf.project(q)).max())
    np.testing.assert_allclose(bf.E, s.K, rtol=1e-14)
    np.testing.assert_allclose(bf.KK, s.KK, rtol=1e-9)   # the oracle has no expm1
    T = np.linspace(1500.0, 4000.0, 7)
    m = s.ctx.mat
    np.testing.assert_allclose(bf.q_evap(T), o.evaporation_flux(
        T, float(m.Pa), float(m.T_boil), float(m.DeltaH_LV), float(m.R_v), float(m.T_liquidus)),
        rtol=1e-12)


@pytest.mark.parametrize("omega", [0.1, 0.5])
def test_converged_step_is_a_fixed_point_of_the_forcing_map(omega):
    s = _after_first_step()
    bf = BruteForce(s.ctx)
    aE = bf.E * s.a
    s.mixing_omega = omega
    _, m = s.step(1e-5, 1e-5)
    assert m["picard_converged"] and m["picard_rel_err"] < 1e-13
    assert m["n_evap_iter"] > 20                      # the nonlinearity takes many iterations
    Ts = bf.surface_T(s.a)
    assert Ts.max() > 3000.0
    q_las = bf.q_laser(1e-5)
    assert bf.q_evap(Ts).max() > 0.1 * q_las.max()    # evaporation is a large term
    Fa = bf.F(aE, s.a, 1e-5)
    inc = np.linalg.norm(s.a - aE)
    assert np.linalg.norm(Fa - s.a) < 1e-10 * inc, np.linalg.norm(Fa - s.a) / inc
    assert np.linalg.norm(Fa - s.a) < 1e-12 * np.linalg.norm(s.a)


def test_fixed_point_does_not_depend_on_the_relaxation():
    ref = _after_first_step()
    aE = BruteForce(ref.ctx).E * ref.a
    runs = {}
    for omega in (0.1, 0.5):
        s = copy.deepcopy(ref)
        s.mixing_omega = omega
        _, m = s.step(1e-5, 1e-5)
        runs[omega] = (s.a.copy(), m["n_evap_iter"])
    (a1, n1), (a5, n5) = runs[0.1], runs[0.5]
    assert n1 != n5
    assert np.linalg.norm(a1 - a5) < 1e-10 * np.linalg.norm(a1 - aE)


def _one_iteration_by_hand(s, omega, cap, shift_sign=1.0):
    """The capped step computed with the brute-force map, from the state of solver ``s``
    (before its step at t = Δt). Returns (a_new, rel, q_evap of the last iterate)."""
    bf = BruteForce(s.ctx)
    dt, t = 1e-5, 1e-5
    st = s.ctx.laser_path.get_state(t, dt)
    aE = bf.E * s.a
    # warm start: previous evaporation translated with the beam, scipy order-1 'constant'
    shift_px = (shift_sign * st.v[1] * dt / bf.dy, shift_sign * st.v[0] * dt / bf.dx)
    q_prev = scipy.ndimage.shift(s.q_evap_old, shift_px, order=1, mode="constant", cval=0.0)
    q_las = bf.q_laser(t)
    a_old = bf.apply(aE, q_las - q_prev)
    N = a_old.size
    for _ in range(cap):
        q_ev = bf.q_evap(bf.surface_T(a_old))
        r = bf.apply(aE, q_las - q_ev) - a_old
        rel = math.sqrt(np.sum(r * r) / N) / max(math.sqrt(np.sum(a_old * a_old) / N), 1e-9)
        a_old = a_old + omega * r
    return a_old, rel, q_ev


@pytest.mark.parametrize("cap", [1, 2])
def test_capped_step_equals_the_relaxed_iteration_by_hand(cap):
    base = _after_first_step()
    assert base.q_evap_old.max() > 0.0
    assert abs(base.ctx.laser_path.get_state(1e-5, 1e-5).v[0]) > 0.0
    s = copy.deepcopy(base)
    s.mixing_omega, s.max_picard_iter, s.convergence_tol = 0.3, cap, 1e-30
    a_hand, rel_hand, q_hand = _one_iteration_by_hand(base, 0.3, cap)
    _, m = s.step(1e-5, 1e-5)
    assert m["n_evap_iter"] == cap and not m["picard_converged"]
    inc = np.linalg.norm(a_hand - BruteForce(base.ctx).E * base.a)
    assert np.linalg.norm(s.a - a_hand) < 1e-12 * inc
    assert m["picard_rel_err"] == pytest.approx(rel_hand, rel=1e-12)
    np.testing.assert_allclose(s.q_evap_old, q_hand, rtol=1e-9, atol=1e-9 * q_hand.max())


def test_warm_start_shift_direction_is_pinned():
    # the beam moves +x at 0.2 cells per step: a flux shifted the other way is a
    # different warm start, and the by-hand iterate then misses the oracle's
    base = _after_first_step()
    s = copy.deepcopy(base)
    s.mixing_omega, s.max_picard_iter, s.convergence_tol = 0.3, 1, 1e-30
    s.step(1e-5, 1e-5)
    good, _, _ = _one_iteration_by_hand(base, 0.3, 1)
    wrong, _, _ = _one_iteration_by_hand(base, 0.3, 1, shift_sign=-1.0)
    inc = np.linalg.norm(good - BruteForce(base.ctx).E * base.a)
    assert np.linalg.norm(s.a - good) < 1e-12 * inc
    assert np.linalg.norm(s.a - wrong) > 1e-4 * inc


# --- the reference's nonlinear invariants, on the oracle --------------------------------

def test_forced_cap_is_reported():
    # tests/integration/test_latent_grid_mode.py:98-110 (grid-mode latent heat, 40×24×60)
    T_S, T_L, T0 = 1674.15, 1697.15, 293.0
    LX, LY, LZ = 0.6e-3, 0.4e-3, 0.15e-3
    cfg = {"simulation": {"duration": 6.0e-5, "dt": 5.0e-6, "max_picard_iter": 3,
                          "dtype": "float64"},
           "domain": {"size": [LX, LY, LZ], "mesh": [40, 24, 60]},
           "material": {"name": "316L", "k": 13.851, "rho": 7957.5, "Cp": 497.89,
                        "T_solidus": T_S, "T_liquidus": T_L, "T0": T0, "L_f": 2.677e5,
                        "h_conv_bottom": 3000.0, "DeltaH_LV": 7.416e6, "R_v": 150.774,
                        "Pa": 101325.0, "T_boil": 3090.0},
           "laser": {"radius": 30e-6, "absorptivity": 0.30, "power_nominal": 120.0}, "io": {}}
    s = o.OracleSolver(make_context(cfg, ConstantVelocityLaser(LX / 2, LY / 2, 0.0, 0.0,
                                                               120.0))).initialize()
    assert s.grid_latent
    ms = [s.step(n * 5e-6, 5e-6)[1] for n in range(2)]
    assert all(m["n_evap_iter"] == 3 for m in ms)
    assert not any(m["picard_converged"] for m in ms)
    assert all(np.isfinite(m["picard_rel_err"]) for m in ms)


@pytest.mark.slow
def test_tdep_box_run_is_physical_with_a_monotone_final_residual():
    # tests/integration/test_property_correction.py:132-170: T-dependent 316L in box mode
    # (refinement 2, 0.4×0.3×0.06 mm box), 48×24×32 modes, 20 steps, cap 80
    k_br = {"solid": "9.248 + 0.01571 * T", "liquid": "12.41 + 0.003279 * T",
            "reference": 24.5075}
    rho_br = {"solid": "8084.2 - 0.42086 * T - 3.8942e-5 * T**2",
              "liquid": "7432.7 + 0.039338 * T - 1.8007e-4 * T**2", "reference": 6896.2365}
    cp_br = {"solid": "458.98 + 0.1328 * T", "liquid": "769.86", "reference": 633.8752}
    T0 = 293.0
    mat = {"name": "316L", "T_solidus": 1674.15, "T_liquidus": 1697.15, "T0": T0,
           "L_f": 2.677e5, "DeltaH_LV": 7.416e6, "R_v": 150.774, "Pa": 101325.0,
           "T_boil": 3090.0, "k": k_br, "rho": rho_br, "Cp": cp_br}
    cfg = {"simulation": {"duration": 1e-3, "dt": 5e-5, "dtype": "float64",
                          "max_picard_iter": 80},
           "domain": {"size": [1e-3, 0.5e-3, 0.25e-3], "mesh": [48, 24, 32]},
           "material": mat,
           "laser": {"radius": 30e-6, "absorptivity": 0.3, "power_nominal": 24.0}, "io": {},
           "fine_mesh": {"refinement": 2, "box_size": [0.4e-3, 0.3e-3, 0.06e-3]}}
    ctx = make_context(cfg, ConstantVelocityLaser(0.25e-3, 0.25e-3, 0.15, 0.0, 24.0))
    s = o.OracleSolver(ctx).initialize()
    assert s.prop_corr and s.fm is not None
    s.track_picard_history = True
    surf_max, its = [], []
    for n in range(20):
        _, m = s.step(n * 5e-5, 5e-5)
        surf_max.append(m["T_surface_max"])
        its.append(m["n_evap_iter"])
    T = o.reconstruct_temperature_DCT(s.a, s.grid)
    assert np.isfinite(T).all()
    assert max(surf_max) < 1.5 * 3090.0
    peak = float(T.max()) - T0
    assert T.min() >= T0 - 0.05 * peak
    Tr = np.linspace(T0, max(surf_max), 200)
    assert s.props.k(Tr).min() > 0.0 and (s.props.rho(Tr) * s.props.c(Tr)).min() > 0.0
    assert its[-1] < 80
    rms = [h["rms_diff"] for h in s.picard_history]
    assert len(rms) == its[-1]
    assert all(b <= a for a, b in zip(rms, rms[1:])), "residual not monotone"


@pytest.mark.slow
def test_full_nonlinear_run_stays_within_physical_bounds():
    # tests/integration/test_validation.py:152-189: latent heat (grid mode) + evaporation,
    # 112×56×112 modes, 40 steps of 25 µs, default Picard controls (ω 0.1, tol 1e-4, cap 30)
    T0 = 293.0
    mat = {"name": "316L", "rho": 7958.0, "k": 13.85, "Cp": 498.0, "L_f": 2.677e5,
           "T_solidus": 1674.15, "T_liquidus": 1697.15, "T0": T0, "DeltaH_LV": 7.416e6,
           "R_v": 150.774, "Pa": 101325.0, "T_boil": 3090.0}
    cfg = {"simulation": {"duration": 1e-3, "dt": 2.5e-5, "dtype": "float64"},
           "domain": {"size": [1e-3, 0.5e-3, 0.25e-3], "mesh": [112, 56, 112]},
           "material": mat,
           "laser": {"radius": 30e-6, "absorptivity": 0.3, "power_nominal": 24.0}, "io": {}}
    s = o.OracleSolver(make_context(cfg, ConstantVelocityLaser(0.25e-3, 0.25e-3, 0.15, 0.0,
                                                               24.0))).initialize()
    assert s.grid_latent and s.max_picard_iter == 30
    surf_max, its = [], []
    for n in range(40):
        _, m = s.step(n * 2.5e-5, 2.5e-5)
        surf_max.append(m["T_surface_max"])
        its.append(m["n_evap_iter"])
    T = o.reconstruct_temperature_DCT(s.a, s.grid)
    assert np.isfinite(T).all()
    assert T.min() >= T0 - 1.0
    assert max(surf_max) < 1.5 * 3090.0
    assert max(surf_max) > 1697.15              # evaporation engages
    assert all(n <= 30 for n in its)
    assert its[-1] < 30

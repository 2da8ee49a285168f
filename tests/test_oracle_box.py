"""Pin the oracle's box mode (laser-following fine sub-box) before it is trusted.

Checks that do not restate the oracle: the reference's own known-answer vectors for the
window centring (tests/unit_tests/physics/test_spectral_helpers.py:32-54), discrete
orthogonality of the fine cosine sums (a full-domain box at any refinement must project
its own reconstruction back to the modes exactly), brute-force modal sums at fine cell
centres, scipy.ndimage.shift for the history shift, the lab-frame consistency of window
slide and history shift, and the reference's integration check that box mode at
refinement 1 over the whole domain reproduces grid mode
(tests/integration/test_latent_grid_mode.py:58-97).
"""

import copy
import math

import numpy as np
import pytest
import scipy.ndimage

from oracle import sgh_oracle as o
from spectral_galerkin_heat_amd.core.laser import ConstantVelocityLaser
from spectral_galerkin_heat_amd.workloads import make_context


@pytest.mark.parametrize("pos,expected", [(5.0, (3, 7, 2)), (0.0, (0, 4, 0)),
                                          (100.0, (16, 20, 3))])
def test_subgrid_indices_reference_vectors(pos, expected):
    assert o.subgrid_indices(pos, 1.0, 20, 4) == expected


def test_subgrid_indices_box_bigger_than_grid():
    start, end, _ = o.subgrid_indices(5.0, 1.0, 20, 30)
    assert (start, end) == (0, 30)


def test_subgrid_indices_round_half_to_even():
    # pos/d = 4.5 rounds to 4 (Python round), so the window starts at 4 - 2 = 2.
    assert o.subgrid_indices(4.5, 1.0, 20, 4)[0] == 2
    assert o.subgrid_indices(5.5, 1.0, 20, 4)[0] == 4


def _fm(size, n, r, box):
    return o.OracleFineMesh(size, n, r, box)


def test_fine_bases_are_brute_force_cosines():
    size, n = (0.6e-3, 0.4e-3, 0.15e-3), (6, 5, 7)
    fm = _fm(size, n, 3, (0.2e-3, 0.25e-3, 0.05e-3))
    fm.update(0.41e-3, 0.07e-3)
    Bx, By, Bz = fm.bases()
    dxf, dyf, dzf = (L / m / 3 for L, m in zip(size, n))
    for a, (B, df, L, m) in enumerate(((Bx, dxf, size[0], n[0]), (By, dyf, size[1], n[1]))):
        for k in range(m):
            for i in range(B.shape[1]):
                x = (fm.start[a] + i + 0.5) * df
                C = math.sqrt((1.0 if k == 0 else 2.0) / L)
                assert B[k, i] == pytest.approx(C * math.cos(math.pi * k * x / L), abs=1e-12)
    # z covers the top box_size_z of the domain: the last fine cell centre is Lz - dzf/2
    z_top = size[2] - 0.5 * dzf
    Cz = math.sqrt(2.0 / size[2])
    assert Bz[1, -1] == pytest.approx(Cz * math.cos(math.pi * z_top / size[2]), abs=1e-12)


@pytest.mark.parametrize("r", [1, 2, 3])
def test_full_domain_box_projects_its_reconstruction_back(r):
    # Discrete orthogonality: Σ_i C_k C_l cos(πk x_i/L) cos(πl x_i/L) d/r = δ_kl for k, l < n·r.
    size, n = (0.5e-3, 0.3e-3, 0.2e-3), (6, 5, 4)
    fm = _fm(size, n, r, size)
    fm.update(0.2e-3, 0.1e-3)
    rng = np.random.default_rng(r)
    a = rng.standard_normal((n[2], n[1], n[0]))
    T = o.reconstruct_box(a, fm)
    assert T.shape == (n[2] * r, n[1] * r, n[0] * r)
    np.testing.assert_allclose(o.project_box(T, fm), a, atol=1e-12)


def test_refinement_one_box_equals_cell_centred_volume():
    size, n = (0.5e-3, 0.3e-3, 0.2e-3), (6, 5, 4)
    fm = _fm(size, n, 1, size)
    g = o.OracleGrid(size, n)
    a = np.random.default_rng(0).standard_normal((4, 5, 6))
    np.testing.assert_allclose(o.reconstruct_box(a, fm), o.reconstruct_volume(a, g), atol=1e-9)
    f = np.random.default_rng(1).standard_normal((4, 5, 6))
    np.testing.assert_allclose(o.project_box(f, fm), o.project_volume(f, g), atol=1e-13)


@pytest.mark.parametrize("nearest,mode", [(True, "nearest"), (False, "constant")])
def test_history_shift_matches_scipy(nearest, mode):
    rng = np.random.default_rng(5)
    f = rng.standard_normal((3, 9, 11))
    for sy, sx in ((0.3, -1.7), (-2.25, 0.6), (0.0, 3.0)):
        ref = scipy.ndimage.shift(f, (0.0, sy, sx), order=1, mode=mode, cval=0.0)
        np.testing.assert_allclose(o.shift_3d_xy(f, sy, sx, nearest), ref, atol=1e-13)


def test_window_slide_and_history_shift_agree_in_the_lab_frame():
    # A field fixed in the lab frame, seen through a window that slides by m fine cells,
    # must be what the shifted history holds (interior columns; the step shifts by
    # -v·dt/d_fine while the window start advances by v·dt/d_fine).
    size, n, r = (0.8e-3, 0.4e-3, 0.2e-3), (16, 8, 4), 2
    fm = _fm(size, n, r, (0.3e-3, 0.4e-3, 0.1e-3))
    dxf = fm.d_fine[0]
    lab = np.random.default_rng(2).standard_normal((fm.n_totals[2], fm.n_totals[1],
                                                    fm.n_totals[0]))
    fm.update(10.0 * dxf, 0.2e-3)
    s0 = fm.start[0]
    seen0 = lab[:, :, s0:s0 + fm.n_box[0]]
    m = 3
    fm.update((10.0 + m) * dxf, 0.2e-3)
    assert fm.start[0] == s0 + m
    seen1 = lab[:, :, fm.start[0]:fm.start[0] + fm.n_box[0]]
    v_dt = m * dxf
    shifted = o.shift_3d_xy(seen0, 0.0, -v_dt / dxf, nearest=True)
    np.testing.assert_allclose(shifted[:, :, :-m], seen1[:, :, :-m], atol=1e-12)


def _latent_cfg(fine=None, maxit=80):
    # tests/integration/test_latent_grid_mode.py:24-41, with a coarser grid
    cfg = {
        "simulation": {"backend": "mi355x", "dtype": "float64", "duration": 6.0e-5,
                       "dt": 5.0e-6, "max_picard_iter": maxit},
        "domain": {"size": [0.6e-3, 0.4e-3, 0.15e-3], "mesh": [20, 12, 30]},
        "material": {"name": "316L", "k": 13.851, "rho": 7957.5, "Cp": 497.89,
                     "T_solidus": 1674.15, "T_liquidus": 1697.15, "T0": 293.0,
                     "L_f": 2.677e5, "h_conv_bottom": 3000.0, "DeltaH_LV": 7.416e6,
                     "R_v": 150.774, "Pa": 101325.0, "T_boil": 3090.0},
        "laser": {"radius": 30e-6, "absorptivity": 0.30, "power_nominal": 120.0},
        "io": {},
    }
    if fine is not None:
        cfg["fine_mesh"] = fine
    return cfg


def _run_oracle(cfg, n):
    ctx = make_context(cfg, ConstantVelocityLaser(0.3e-3, 0.2e-3, 0.0, 0.0, 120.0))
    s = o.OracleSolver(ctx).initialize()
    mets = [s.step(k * ctx.num.dt, ctx.num.dt)[1] for k in range(n)]
    return s, mets


def test_box_refinement_one_reproduces_grid_mode():
    # tests/integration/test_latent_grid_mode.py:58-97: same resolution, same answer, and
    # both modes warm-start from the previous latent source (equal iteration counts).
    sg, mg = _run_oracle(_latent_cfg(), 6)
    sb, mb = _run_oracle(_latent_cfg({"refinement": 1, "box_size": [0.6e-3, 0.4e-3, 0.15e-3]}),
                         6)
    assert sg.fm is None and sg.grid_latent
    assert sb.fm is not None and not sb.grid_latent
    assert all(m["picard_converged"] for m in mg + mb)
    assert [m["n_evap_iter"] for m in mg] == [m["n_evap_iter"] for m in mb]
    assert o.reconstruct_volume(sg.a, sg.grid).max() > 1697.15
    np.testing.assert_allclose(sb.a, sg.a, rtol=0, atol=1e-10 * np.abs(sg.a).max())


def test_box_mode_latent_heat_lowers_the_peak():
    # With L_f > 0 the melting front absorbs energy, so the peak temperature of a box-mode
    # run must lie below the same run with L_f = 0 (the sign of the latent sink).
    cfg = _latent_cfg({"refinement": 2, "box_size": [0.3e-3, 0.3e-3, 0.05e-3]})
    s1, _ = _run_oracle(cfg, 6)
    cfg0 = copy.deepcopy(cfg)
    cfg0["material"]["L_f"] = 0.0
    s0, _ = _run_oracle(cfg0, 6)
    assert o.reconstruct_volume(s1.a, s1.grid).max() < o.reconstruct_volume(s0.a, s0.grid).max() - 1.0
    assert s1.fm.Q_prev.min() < 0.0

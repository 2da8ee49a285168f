"""GPU parity of box mode (the laser-following fine sub-box) against the FP64 oracle.

Kernel level: the strided MFMA contraction against einsum, the box reconstruction and
projection against the oracle's einsums on sliding windows, the history shift against
the oracle (which is pinned to scipy.ndimage.shift), and the face peaks against NumPy.
Step level: SpectralSolver on the mi355x backend and the oracle side by side, FP64,
relative L2 ≤ 1e-9 of the modal state and of the box history with identical Picard
counts, on (a) the reference's box-equals-grid configuration (refinement 1, full
domain), (b) a moving track with sliding x and y windows at refinement 2, (c) the
temperature-dependent case with property correction.
"""

import copy
import logging

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import sgh_oracle as o  # noqa: E402
from spectral_galerkin_heat_amd import _lib  # noqa: E402
from spectral_galerkin_heat_amd.core.laser import ConstantVelocityLaser  # noqa: E402
from spectral_galerkin_heat_amd.workloads import (make_context, single_track_box,  # noqa: E402
                                                  tdep_track_box)

DEV = "cuda"


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV, dtype)


def host(t):
    return t.detach().cpu().double().numpy()


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape_in,n_out,ldm_extra", [((3, 9, 11), 14, 5), ((2, 17, 1), 7, 30),
                                                      ((1, 24, 300), 40, 1)])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-6)])
def test_strided_contraction_reads_a_column_window(shape_in, n_out, ldm_extra, dtype, tol):
    rng = np.random.default_rng(sum(shape_in) + n_out)
    outer, n_in, inner = shape_in
    off = 3
    wide = rng.standard_normal((n_out, off + n_in + ldm_extra))
    M = wide[:, off:off + n_in]
    X = rng.standard_normal(shape_in)
    Wt = dev(wide, dtype)
    Xt = dev(X, dtype)
    out = torch.empty((outer, n_out, inner), dtype=dtype, device=DEV)
    _lib.call("sgh_axis_contract_strided", _lib.dtype_code(dtype), 0, outer, n_in, n_out, inner,
              Wt.data_ptr() + off * Wt.element_size(), wide.shape[1], _lib.ptr(Xt), 0.5,
              _lib.ptr(out), _lib.stream_ptr())
    ref = 0.5 * np.einsum("kj,oji->oki", M, X)
    assert rel(host(out), ref) < tol


def test_strided_contraction_rejects_short_stride():
    M = dev(np.eye(4))
    X = dev(np.ones((2, 4, 1)))
    out = torch.empty_like(X)
    with pytest.raises(_lib.SGHError, match="ldm"):
        _lib.call("sgh_axis_contract_strided", _lib.F64, 0, 2, 4, 4, 1, _lib.ptr(M), 3,
                  _lib.ptr(X), 1.0, _lib.ptr(out), _lib.stream_ptr())


def _box_state(size, mesh, r, box, dtype=np.float64):
    from spectral_galerkin_heat_amd.core.parameters import (FineMeshParams, GeomParams,
                                                            MaterialParams, NumParams)
    from spectral_galerkin_heat_amd.core.vector import Vec3
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as K
    geom = GeomParams(size=Vec3(*size), n=Vec3(*mesh))
    num = NumParams(dt=1e-6, nx=mesh[0], ny=mesh[1], nz=mesh[2], dtype=dtype)
    fine = FineMeshParams(refinement=r, box_size=Vec3(*box))
    S = K.SpectralSolverState(MaterialParams(rho=7900.0, k=20.0, Cp=500.0), geom, num, fine)
    return K, S, o.OracleFineMesh(size, mesh, r, box)


class _At:
    def __init__(self, x, y):
        self.x, self.y, self.v = x, y, (0.0, 0.0)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
@pytest.mark.parametrize("pos", [(0.05e-3, 0.35e-3), (0.31e-3, 0.2e-3), (0.59e-3, 0.01e-3)])
def test_box_reconstruction_and_projection_match_oracle(dtype, tol, pos):
    size, mesh = (0.6e-3, 0.4e-3, 0.15e-3), (20, 12, 30)
    K, S, fm = _box_state(size, mesh, 2, (0.25e-3, 0.15e-3, 0.05e-3), dtype)
    S.fine_mesh.update(_At(*pos))
    fm.update(*pos)
    assert [S.fine_mesh.x_start, S.fine_mesh.y_start] == fm.start[:2]
    assert list(S.fine_mesh.shape) == fm.n_box[::-1]
    rng = np.random.default_rng(4)
    a = rng.standard_normal((30, 12, 20))
    a[0, 0, 0] += 50.0
    td = torch.float64 if dtype == np.float64 else torch.float32
    T = K._reconstruct_temperature_box(dev(a, td), S)
    assert rel(host(T), o.reconstruct_box(a, fm)) < tol
    f = rng.standard_normal(fm.n_box[::-1])
    m = K.project_box_to_modes(dev(f, td), S)
    assert rel(host(m), o.project_box(f, fm)) < tol


@pytest.mark.parametrize("dtype,variant,tol", [(np.float64, 0, 1e-11), (np.float64, 2, 1e-11),
                                               (np.float32, 0, 2e-4), (np.float32, 1, 5e-3)])
@pytest.mark.parametrize("tdep", [False, True])
def test_fused_box_latent_source_matches_oracle(dtype, variant, tol, tdep):
    from spectral_galerkin_heat_amd.core.properties import MaterialModel
    size, mesh = (0.6e-3, 0.4e-3, 0.15e-3), (20, 12, 30)
    K, S, fm = _box_state(size, mesh, 2, (0.25e-3, 0.15e-3, 0.05e-3), dtype)
    S.xform = variant
    pos = (0.31e-3, 0.12e-3)
    S.fine_mesh.update(_At(*pos))
    fm.update(*pos)
    rng = np.random.default_rng(12)
    a = rng.standard_normal((30, 12, 20))
    a *= 20.0 / o.reconstruct_box(a, fm).std()
    a[0, 0, 0] += 1685.0 * np.sqrt(np.prod(size))
    Tp = o.reconstruct_box(a, fm) + rng.uniform(-15.0, 15.0, fm.n_box[::-1])
    td = torch.float64 if dtype == np.float64 else torch.float32
    S.fine_mesh.T_prev = dev(Tp, td)

    class Phys:
        T_solidus, T_liquidus, rho, L_f = 1674.15, 1697.15, 7957.5, 2.677e5
        model = None

    class Num:
        dt = 5e-6

    T = o.reconstruct_box(a, fm)
    if tdep:
        Phys.model = MaterialModel.from_config({
            "k": {"solid": "9.248 + 0.01571 * T", "liquid": "12.41 + 0.003279 * T",
                  "reference": 24.5},
            "rho": {"solid": "8084.2 - 0.42086 * T - 3.8942e-5 * T**2",
                    "liquid": "7432.7 + 0.039338 * T - 1.8007e-4 * T**2", "reference": 6896.0},
            "Cp": {"solid": "458.98 + 0.1328 * T", "liquid": "769.86", "reference": 633.9}},
            1674.15, 1697.15)
        ref = o.OracleProps.from_model(Phys.model).latent(T, Tp, Phys.L_f, Num.dt)
    else:
        ref = o.latent_source(T, Tp, 1674.15, 1697.15, 7957.5, 2.677e5, 5e-6)
    assert np.abs(ref).max() > 0
    Q = torch.empty(S.fine_mesh.shape, dtype=td, device=DEV)
    K._box_latent(Q, dev(a, td), S, Phys, Num)
    assert rel(host(Q), ref) < tol


@pytest.mark.parametrize("nearest", [1, 0])
def test_history_shift_matches_oracle(nearest):
    rng = np.random.default_rng(6)
    f = rng.standard_normal((5, 17, 23))
    src = dev(f)
    out = torch.empty_like(src)
    for sy, sx in ((0.3, -1.7), (-2.25, 0.6), (0.0, 3.0), (0.0, -0.0)):
        _lib.call("sgh_shift3d_xy", _lib.F64, 5, 17, 23, _lib.ptr(src), sy, sx, nearest,
                  _lib.ptr(out), _lib.stream_ptr())
        ref = o.shift_3d_xy(f, sy, sx, bool(nearest))
        np.testing.assert_allclose(host(out), ref, atol=1e-14)


def test_face_peaks_match_numpy():
    rng = np.random.default_rng(8)
    Q = rng.standard_normal((6, 9, 13))
    Q[2, 4, 6] = -40.0
    out = torch.zeros(4 * 7, dtype=torch.float64, device=DEV)
    _lib.call("sgh_box_face_peaks", _lib.F64, 6, 9, 13, _lib.ptr(dev(Q)), _lib.ptr(out),
              _lib.stream_ptr())
    A = np.abs(Q)
    ref = [A.max(), A[0].max(), A[:, :, (0, -1)].max(), A[:, (0, -1), :].max()]
    np.testing.assert_allclose(host(out)[:4], ref, rtol=0, atol=0)


# --- step level -----------------------------------------------------------------------

def _solver(cfg, lp):
    from spectral_galerkin_heat_amd.backends import get_backend
    from spectral_galerkin_heat_amd.solvers.spectral import SpectralSolver
    ctx = make_context(cfg, lp)
    s = SpectralSolver(get_backend("mi355x"), ctx)
    s.initialize()
    return s, ctx


def _run_pair(cfg, lp_factory, n_steps):
    s, ctx = _solver(cfg, lp_factory())
    ref = o.OracleSolver(make_context(cfg, lp_factory())).initialize()
    dt = ctx.num.dt
    its, its_ref = [], []
    for n in range(n_steps):
        _, m = s.step(n * dt, dt)
        _, mr = ref.step(n * dt, dt)
        its.append(m["n_evap_iter"])
        its_ref.append(mr["n_evap_iter"])
    return s, ref, its, its_ref, m, mr


def _latent_cfg(fine):
    cfg = {
        "simulation": {"backend": "mi355x", "dtype": "float64", "duration": 6.0e-5,
                       "dt": 5.0e-6, "max_picard_iter": 80},
        "domain": {"size": [0.6e-3, 0.4e-3, 0.15e-3], "mesh": [20, 12, 30]},
        "material": {"name": "316L", "k": 13.851, "rho": 7957.5, "Cp": 497.89,
                     "T_solidus": 1674.15, "T_liquidus": 1697.15, "T0": 293.0,
                     "L_f": 2.677e5, "h_conv_bottom": 3000.0, "DeltaH_LV": 7.416e6,
                     "R_v": 150.774, "Pa": 101325.0, "T_boil": 3090.0},
        "laser": {"radius": 30e-6, "absorptivity": 0.30, "power_nominal": 120.0},
        "io": {},
        "fine_mesh": fine,
    }
    return cfg


def test_box_refinement_one_matches_oracle_and_grid_mode():
    cfg = _latent_cfg({"refinement": 1, "box_size": [0.6e-3, 0.4e-3, 0.15e-3]})

    def lp():
        return ConstantVelocityLaser(0.3e-3, 0.2e-3, 0.0, 0.0, 120.0)

    s, ref, its, its_ref, m, mr = _run_pair(cfg, lp, 6)
    assert s.state.fine_mesh is not None and not s._grid_latent
    assert its == its_ref
    assert rel(host(s.state.a), ref.a) < 1e-9
    assert rel(host(s.state.fine_mesh.T_prev), ref.fm.T_prev) < 1e-9
    assert rel(host(s.state.fine_mesh.Q_prev), ref.fm.Q_prev) < 1e-8
    # and the same as grid mode on the device (tests/integration/test_latent_grid_mode.py:58-97)
    g_cfg = copy.deepcopy(cfg)
    del g_cfg["fine_mesh"]
    sg, ctx = _solver(g_cfg, lp())
    its_g = [sg.step(n * ctx.num.dt, ctx.num.dt)[1]["n_evap_iter"] for n in range(6)]
    assert sg._grid_latent and its_g == its
    assert rel(host(sg.state.a), host(s.state.a)) < 1e-9


def test_moving_track_sliding_windows_fp64_matches_oracle_1e9():

    def lp():
        return single_track_box(mesh=(40, 24, 60), box_size=(0.3e-3, 0.3e-3, 0.05e-3))[1]

    cfg = single_track_box(mesh=(40, 24, 60), box_size=(0.3e-3, 0.3e-3, 0.05e-3))[0]
    s, ref, its, its_ref, m, mr = _run_pair(cfg, lp, 10)
    fm = s.state.fine_mesh
    assert fm.nx_box < fm.n_fine_totals[0] and fm.ny_box < fm.n_fine_totals[1]
    assert [fm.x_start, fm.y_start] == ref.fm.start[:2] and fm.x_start > 0
    assert its == its_ref
    assert rel(host(s.state.a), ref.a) < 1e-9
    assert rel(host(fm.T_prev), ref.fm.T_prev) < 1e-9
    assert rel(host(fm.Q_prev), ref.fm.Q_prev) < 1e-8
    assert float(m["T_surface_max"]) == pytest.approx(mr["T_surface_max"], rel=1e-9)


def test_tdep_box_with_property_correction_fp64_matches_oracle_1e9():
    def lp():
        return tdep_track_box(mesh=(50, 30, 48))[1]

    cfg = tdep_track_box(mesh=(50, 30, 48))[0]
    s, ref, its, its_ref, m, mr = _run_pair(cfg, lp, 3)
    assert s._property_correction and s.state.fine_mesh is not None
    assert its == its_ref
    assert rel(host(s.state.a), ref.a) < 1e-9
    assert rel(host(s._T_prev_full), ref.T_prev_full) < 1e-9
    assert rel(host(s.state.fine_mesh.T_prev), ref.fm.T_prev) < 1e-9


def test_shallow_box_warns_once_like_the_oracle(caplog):
    # A box one fine layer deep: the melt front reaches its floor, as the oracle reports.
    cfg = _latent_cfg({"refinement": 2, "box_size": [0.3e-3, 0.3e-3, 2.5e-6]})

    def lp():
        return ConstantVelocityLaser(0.3e-3, 0.2e-3, 0.0, 0.0, 120.0)

    with caplog.at_level(logging.WARNING):
        s, ref, its, its_ref, m, mr = _run_pair(cfg, lp, 6)
    assert its == its_ref
    assert ref.truncation_faces and "floor (depth)" in ref.truncation_faces
    warned = [r for r in caplog.records if "fine-mesh box boundary" in r.getMessage()]
    assert len(warned) == 1 and "floor (depth)" in warned[0].getMessage()
    assert s._latent_truncation_warned


@pytest.mark.parametrize("mesh,r,box,start,vel", [
    ((21, 15, 17), 3, (0.2e-3, 0.15e-3, 0.04e-3), (0.02e-3, 0.37e-3), (0.0, -0.3)),
    ((16, 20, 24), 2, (0.45e-3, 0.1e-3, 0.06e-3), (0.58e-3, 0.05e-3), (-0.5, 0.2)),
    ((25, 13, 30), 1, (0.3e-3, 0.25e-3, 0.15e-3), (0.3e-3, 0.2e-3), (0.4, 0.4)),
])
def test_box_odd_sizes_edges_and_diagonal_motion_fp64_matches_oracle(mesh, r, box, start, vel):
    # odd mode counts (rows not a multiple of 16 bytes: register staging paths), windows
    # pinned at both domain edges, motion along y and diagonally, refinement 1 to 3
    cfg = _latent_cfg({"refinement": r, "box_size": list(box)})
    cfg["domain"]["mesh"] = list(mesh)

    def lp():
        return ConstantVelocityLaser(start[0], start[1], vel[0], vel[1], 120.0)

    s, ref, its, its_ref, m, mr = _run_pair(cfg, lp, 5)
    assert its == its_ref
    assert rel(host(s.state.a), ref.a) < 1e-9
    assert rel(host(s.state.fine_mesh.T_prev), ref.fm.T_prev) < 1e-9
    assert [s.state.fine_mesh.x_start, s.state.fine_mesh.y_start] == ref.fm.start[:2]


def test_box_fp32_within_1e4_of_the_fp64_oracle():
    cfg = _latent_cfg({"refinement": 2, "box_size": [0.3e-3, 0.3e-3, 0.05e-3]})
    cfg["simulation"]["dtype"] = "float32"

    def lp():
        return ConstantVelocityLaser(0.2e-3, 0.2e-3, 0.4, 0.0, 120.0)

    s, ctx = _solver(cfg, lp())
    cfg64 = copy.deepcopy(cfg)
    cfg64["simulation"]["dtype"] = "float64"
    ref = o.OracleSolver(make_context(cfg64, lp())).initialize()
    dt = ctx.num.dt
    for n in range(6):
        s.step(n * dt, dt)
        ref.step(n * dt, dt)
    assert rel(host(s.state.a), ref.a) < 1e-4

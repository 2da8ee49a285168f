"""GPU checks of the 3-D axis contractions (``sgh_axis_contract``) in every arithmetic.

Each variant contracts one axis of a seeded [outer, n, inner] volume with an n×n matrix
and is compared with the FP64 einsum of the same operands. Tolerances (relative L2):
MFMA f64 and vector f64 1e-13; MFMA f32 and vector f32 2e-6; split-BF16 (three bf16
products, FP32 accumulation) 3e-5. Shapes include n and inner that are not multiples of
the 64×64×32 tile, a batch of planes, and outer large enough for several row tiles.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import sgh_oracle as o  # noqa: E402
from spectral_galerkin_heat_amd import _lib  # noqa: E402

DEV = "cuda"
MFMA, BF16X3, VALU = 0, 1, 2
TOL = {(MFMA, torch.float64): 1e-13, (VALU, torch.float64): 1e-13,
       (MFMA, torch.float32): 2e-6, (VALU, torch.float32): 2e-6,
       (BF16X3, torch.float32): 3e-5}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def contract(variant, dtype, M, X, scale=1.0):
    outer, n, inner = X.shape
    Mt = torch.as_tensor(M).to(DEV, dtype).contiguous()
    Xt = torch.as_tensor(X).to(DEV, dtype).contiguous()
    out = torch.empty_like(Xt)
    _lib.call("sgh_axis_contract", _lib.dtype_code(dtype), variant, outer, n, inner, _lib.ptr(Mt),
              _lib.ptr(Xt), float(scale), _lib.ptr(out), _lib.stream_ptr())
    torch.cuda.synchronize()
    return out.cpu().double().numpy()


@pytest.mark.parametrize("variant,dtype", sorted(TOL, key=str))
@pytest.mark.parametrize("shape", [(5, 7, 1), (3, 9, 11), (1, 24, 4096), (16, 48, 96),
                                   (130, 70, 1), (2, 130, 70), (64, 128, 1), (1, 32, 8192)])
def test_axis_contract_matches_fp64_einsum(variant, dtype, shape):
    rng = np.random.default_rng(sum(shape))
    outer, n, inner = shape
    M = rng.standard_normal((n, n))
    X = rng.standard_normal(shape)
    X[..., 0, :] += 300.0   # a dominant mean mode, as in a temperature field
    ref = 0.5 * np.einsum("kj,oji->oki", M, X)
    got = contract(variant, dtype, M, X, scale=0.5)
    assert rel(got, ref) < TOL[(variant, dtype)]


@pytest.mark.parametrize("shape", [(12000, 160, 1), (22050, 192, 1), (4100, 80, 1),
                                   (37, 96, 160), (112, 96, 160), (301, 128, 192)])
def test_axis_contract_tile_shapes_match_fp64_einsum(shape):
    # sizes on both sides of the 32-row / 64-row output-tile choice (below 256 workgroups of
    # 64-row tiles in FP64, 1024 in FP32), in both layouts (inner = 1: rows of the volume;
    # inner > 1: planes of a batch), with tile counts that are and are not multiples of the
    # 8 XCDs the tiles are dealt over
    rng = np.random.default_rng(sum(shape))
    outer, n, inner = shape
    M = rng.standard_normal((n, n))
    X = rng.standard_normal(shape)
    X[..., 0, :] += 300.0
    ref = 0.5 * np.einsum("kj,oji->oki", M, X, optimize=True)
    for variant, dtype in ((MFMA, torch.float64), (MFMA, torch.float32),
                           (BF16X3, torch.float32)):
        got = contract(variant, dtype, M, X, scale=0.5)
        assert rel(got, ref) < TOL[(variant, dtype)], (variant, dtype)


@pytest.mark.parametrize("shape", [(16, 32, 64), (32, 64, 128), (24, 48, 96)])
def test_dct3d_every_variant_matches_scipy(shape):
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as K
    rng = np.random.default_rng(3)
    f = 293.0 + 50.0 * rng.standard_normal(shape)
    ref = o.dct2(f)
    for (variant, dtype), tol in TOL.items():
        x = torch.as_tensor(f).to(DEV, dtype)
        mats = [K._dct_mats(n, dtype)[0] for n in (shape[2], shape[1], shape[0])]
        got = K._apply_axes(x, mats, xform=variant).cpu().double().numpy()
        assert rel(got, ref) < 3 * tol, (variant, dtype)


def test_bf16x3_rejects_fp64_data():
    M = np.eye(8)
    X = np.ones((2, 8, 3))
    with pytest.raises(_lib.SGHError, match="split-BF16"):
        contract(BF16X3, torch.float64, M, X)


def test_solver_state_selects_transform():
    from spectral_galerkin_heat_amd.core.parameters import GeomParams, MaterialParams, NumParams
    from spectral_galerkin_heat_amd.core.vector import Vec3
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as K
    geom = GeomParams(size=Vec3(1e-3, 1e-3, 5e-4), n=Vec3(16, 16, 8))
    phys = MaterialParams(rho=7900.0, k=20.0, Cp=500.0)
    num = NumParams(dt=1e-6, nx=16, ny=16, nz=8, dtype=np.float32, transform="bf16x3")
    S = K.SpectralSolverState(phys, geom, num, None)
    assert S.xform == BF16X3
    a = torch.zeros((8, 16, 16), device=DEV)
    a[0, 0, 0] = 293.0 * np.sqrt(1e-3 * 1e-3 * 5e-4)
    T = K.reconstruct_volume(a, S)
    np.testing.assert_allclose(T.cpu().numpy(), 293.0, rtol=2e-5)


@pytest.mark.parametrize("shape_in,n_out", [((3, 9, 11), 10), ((2, 17, 1), 18), ((1, 24, 4096), 25)])
def test_rectangular_axis_contract(shape_in, n_out):
    rng = np.random.default_rng(7)
    outer, n_in, inner = shape_in
    M = rng.standard_normal((n_out, n_in))
    X = rng.standard_normal(shape_in)
    Mt = torch.as_tensor(M).to(DEV).contiguous()
    Xt = torch.as_tensor(X).to(DEV).contiguous()
    out = torch.empty((outer, n_out, inner), dtype=torch.float64, device=DEV)
    _lib.call("sgh_axis_contract_rect", _lib.F64, MFMA, outer, n_in, n_out, inner, _lib.ptr(Mt),
              _lib.ptr(Xt), 2.0, _lib.ptr(out), _lib.stream_ptr())
    ref = 2.0 * np.einsum("kj,oji->oki", M, X)
    assert rel(out.cpu().numpy(), ref) < 1e-13


@pytest.mark.parametrize("mesh,dtype,tol", [((12, 10, 8), torch.float64, 1e-13),
                                            ((64, 32, 16), torch.float64, 1e-13),
                                            ((48, 40, 24), torch.float32, 2e-6)])
def test_node_reconstruction_matches_oracle_dct1(mesh, dtype, tol):
    """reconstruct_temperature_DCT on the device vs the oracle's padded DCT-I."""
    from spectral_galerkin_heat_amd.core.parameters import GeomParams, MaterialParams, NumParams
    from spectral_galerkin_heat_amd.core.vector import Vec3
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as K
    from spectral_galerkin_heat_amd.physics.spectral_helpers import reconstruct_temperature_DCT
    nx, ny, nz = mesh
    size = (1.3e-3, 0.9e-3, 0.4e-3)
    geom = GeomParams(size=Vec3(*size), n=Vec3(*mesh))
    num = NumParams(dt=1e-6, nx=nx, ny=ny, nz=nz,
                    dtype=np.float64 if dtype == torch.float64 else np.float32)
    S = K.SpectralSolverState(MaterialParams(rho=7900.0, k=20.0, Cp=500.0), geom, num, None)
    rng = np.random.default_rng(11)
    a = rng.standard_normal((nz, ny, nx))
    a[0, 0, 0] += 300.0
    got = reconstruct_temperature_DCT(torch.as_tensor(a).to(DEV, dtype), S)
    assert tuple(got.shape) == (nx + 1, ny + 1, nz + 1)
    ref = o.reconstruct_temperature_DCT(a, o.OracleGrid(size, mesh))
    assert rel(got.cpu().double().numpy(), ref) < tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_point_evaluation_matches_oracle(dtype, tol):
    from spectral_galerkin_heat_amd.core.parameters import GeomParams, MaterialParams, NumParams
    from spectral_galerkin_heat_amd.core.vector import Vec3
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as K
    from spectral_galerkin_heat_amd.physics.spectral_helpers import (
        reconstruct_temperature_volume_at_points)
    mesh, size = (24, 16, 12), (1.2e-3, 0.8e-3, 0.3e-3)
    geom = GeomParams(size=Vec3(*size), n=Vec3(*mesh))
    num = NumParams(dt=1e-6, nx=mesh[0], ny=mesh[1], nz=mesh[2],
                    dtype=np.float64 if dtype == torch.float64 else np.float32)
    S = K.SpectralSolverState(MaterialParams(rho=7900.0, k=20.0, Cp=500.0), geom, num, None)
    rng = np.random.default_rng(21)
    a = rng.standard_normal((12, 16, 24))
    a[0, 0, 0] += 300.0
    g = o.OracleGrid(size, mesh)
    pts = rng.uniform(-0.1e-3, 1.3e-3, (257, 3)) * np.array([1.0, 0.7, 0.25])
    got = reconstruct_temperature_volume_at_points(torch.as_tensor(a).to(DEV, dtype), num, geom,
                                                   S, pts).cpu().double().numpy()
    assert rel(got, o.reconstruct_at_points(a, g, pts)) < tol
    # at the cell centres the series is the cell-centred volume reconstruction
    zc, yc, xc = np.meshgrid(g.z, g.y, g.x, indexing="ij")
    centres = np.stack([xc.ravel(), yc.ravel(), zc.ravel()], axis=1)
    got_c = reconstruct_temperature_volume_at_points(torch.as_tensor(a).to(DEV, dtype), num,
                                                     geom, S, centres).cpu().double().numpy()
    assert rel(got_c, o.reconstruct_volume(a, g).ravel()) < tol

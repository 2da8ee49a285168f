"""The fused Picard pass with its norm reduction folded in (last workgroup sums the
per-block pairs) against the two-launch form: identical iterate, partials and norms, bit
for bit, on two launches in a row (the completion counter is left at zero); the copy the
pass writes to pinned host memory carries the same norms and the sequence number."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spectral_galerkin_heat_amd import _lib  # noqa: E402

DEV = "cuda"


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape,nchunks", [((16, 12, 20), 8), ((30, 48, 80), 8), ((3, 5, 7), 2)])
def test_fused_norm_pass_equals_separate_reduction(dtype, shape, nchunks):
    nz, ny, nx = shape
    rng = np.random.default_rng(nz * ny * nx)
    t = lambda *s: torch.as_tensor(rng.standard_normal(s)).to(DEV, dtype)  # noqa: E731
    aK, KK, a_old, Q = t(nz, ny, nx), t(nz, ny, nx), t(nz, ny, nx), t(nz, ny, nx)
    wt, St, wz = t(nz), t(ny, nx), t(nz)
    code = _lib.dtype_code(dtype)
    n_pairs = _lib.lib.sgh_picard_partials_count(nx, ny, nz, nchunks)
    st = _lib.stream_ptr()
    outs = []
    hp = ctypes.c_void_p()
    _lib.call("sgh_host_alloc", 64, ctypes.byref(hp))
    host = np.ctypeslib.as_array((ctypes.c_double * 8).from_address(hp.value))
    host[:] = -1.0
    for n, fused in enumerate((False, True, True)):
        a_new = torch.empty_like(a_old)
        zc = torch.zeros((nchunks, ny, nx), dtype=dtype, device=DEV)
        part = torch.zeros(2 * n_pairs, dtype=torch.float64, device=DEV)
        norms = torch.zeros(3, dtype=torch.float64, device=DEV)
        args = (code, nx, ny, nz, _lib.ptr(aK), _lib.ptr(KK), _lib.ptr(a_old), _lib.ptr(wt),
                _lib.ptr(St), None, None, _lib.ptr(Q), 0.3, _lib.ptr(wz), nchunks, _lib.ptr(zc),
                _lib.ptr(a_new), _lib.ptr(part))
        if fused:
            if not outs[1:]:
                counter = torch.zeros(1, dtype=torch.int32, device=DEV)
            _lib.call("sgh_picard_iter_fused_norms", *args, _lib.ptr(counter), _lib.ptr(norms),
                      hp.value + 32 * (n - 1), float(10 + n), st)
        else:
            _lib.call("sgh_picard_iter_fused", *args, st)
            _lib.call("sgh_reduce_norms", _lib.ptr(part), n_pairs, float(nz * ny * nx),
                      _lib.ptr(norms), st)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in (a_new, zc, part, norms)])
    assert int(counter.item()) == 0
    for got in outs[1:]:
        for g, r in zip(got, outs[0]):
            np.testing.assert_array_equal(g, r)
    for n in (1, 2):
        np.testing.assert_array_equal(host[4 * (n - 1):4 * (n - 1) + 3], outs[0][3])
        assert host[4 * (n - 1) + 3] == 10 + n
    _lib.call("sgh_host_free", hp.value)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("ny,nx,tdep", [(48, 80, False), (32, 64, True), (64, 128, True),
                                       (20, 25, False), (128, 256, False)])
def test_fused_inverse_transform_and_forcing_equal_the_two_calls(dtype, ny, nx, tdep):
    from spectral_galerkin_heat_amd.core.properties import MaterialModel
    from spectral_galerkin_heat_amd.physics import mi355x_kernels as K
    rng = np.random.default_rng(ny + nx)
    L = torch.as_tensor(K.dct_ortho_matrix(ny).T.copy()).to(DEV, dtype)
    R = torch.as_tensor(K.dct_ortho_matrix(nx)).to(DEV, dtype)
    X = torch.as_tensor(rng.standard_normal((ny, nx)) * 400.0).to(DEV, dtype)
    X[0, 0] += 2000.0 * np.sqrt(nx * ny)
    q_las = torch.as_tensor(rng.uniform(0, 1e9, (ny, nx))).to(DEV, dtype)
    prog = None
    if tdep:
        model = MaterialModel.from_config({
            "k": {"solid": "9.248 + 0.01571 * T", "liquid": "12.41 + 0.003279 * T",
                  "reference": 24.5},
            "rho": {"solid": "7900.0", "liquid": "7400.0", "reference": 7650.0},
            "Cp": {"solid": "458.98 + 0.1328 * T", "liquid": "769.86", "reference": 633.9}},
            1674.15, 1697.15)
        prog = K._program(model)
    code, st = _lib.dtype_code(dtype), _lib.stream_ptr()
    phys = (101325.0, 3090.0, 7.416e6, 150.774, 1697.15, 10.0, 293.0)
    T_ref = torch.empty_like(X)
    qe_ref, qn_ref = torch.empty_like(X), torch.empty_like(X)
    _lib.call("sgh_sep2d", code, 1, ny, nx, _lib.ptr(L), _lib.ptr(R), _lib.ptr(X), 1, 1.0,
              _lib.ptr(T_ref), st)
    _lib.call("sgh_surface_forcing", code, nx * ny, _lib.ptr(T_ref), _lib.ptr(q_las), *phys,
              _lib.ptr(prog), 24.5 if tdep else 0.0, _lib.ptr(qe_ref), _lib.ptr(qn_ref), st)
    T_s = torch.zeros_like(X)
    qe, qn = torch.empty_like(X), torch.empty_like(X)
    _lib.call("sgh_sep2d_forcing", code, ny, nx, _lib.ptr(L), _lib.ptr(R), _lib.ptr(X),
              1.0, _lib.ptr(q_las), *phys, _lib.ptr(prog),
              24.5 if tdep else 0.0, _lib.ptr(T_s), _lib.ptr(qe), _lib.ptr(qn), st)
    torch.cuda.synchronize()
    assert float(qe_ref.abs().max()) > 0.0    # some cells are above the liquidus
    np.testing.assert_array_equal(qe.cpu().numpy(), qe_ref.cpu().numpy())
    np.testing.assert_array_equal(qn.cpu().numpy(), qn_ref.cpu().numpy())

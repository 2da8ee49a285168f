"""The allocation-free fused Picard pass (sgh_picard_iter_fused_norms_noalloc) against the
entry point with caller scratch (sgh_picard_iter_fused_norms): identical bits for the new
iterate, the next surface z-contraction and the three norms, on two calls in a row (the
second smaller, so the fixed scratch keeps the first call's pairs beyond its range), in
FP64 and FP32, through the C ABI."""

import ctypes as C

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(nx, ny, nz, dtype, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)

    def r(*shape, scale=1.0):
        return (torch.randn(*shape, generator=g, dtype=torch.float64) * scale).to(
            device="cuda", dtype=dtype).contiguous()
    N = (nz, ny, nx)
    return {"aK": r(*N, scale=300.0), "KK": r(*N, scale=1e-9), "a_old": r(*N, scale=300.0),
            "wt": r(nz), "St": r(ny, nx, scale=1e7), "wb": r(nz), "Sb": r(ny, nx, scale=1e6),
            "Q": r(*N, scale=1e5), "wz": r(nz)}


def _run(lib, fn, t, nx, ny, nz, nchunks, code, dtype, omega, scratch=None):
    a_new = torch.empty((nz, ny, nx), dtype=dtype, device="cuda")
    zc = torch.empty((nchunks, ny, nx), dtype=dtype, device="cuda")
    norms = torch.zeros(3, dtype=torch.float64, device="cuda")
    p = [t[k].data_ptr() for k in ("aK", "KK", "a_old", "wt", "St", "wb", "Sb", "Q")]
    s = torch.cuda.current_stream().cuda_stream
    if scratch is None:
        rc = lib.sgh_picard_iter_fused_norms_noalloc(
            code, nx, ny, nz, *p, omega, t["wz"].data_ptr(), nchunks, zc.data_ptr(),
            a_new.data_ptr(), C.cast(norms.data_ptr(), C.POINTER(C.c_double)), None, 0.0, s)
    else:
        partials, counter = scratch
        rc = lib.sgh_picard_iter_fused_norms(
            code, nx, ny, nz, *p, omega, t["wz"].data_ptr(), nchunks, zc.data_ptr(),
            a_new.data_ptr(), C.cast(partials.data_ptr(), C.POINTER(C.c_double)),
            C.cast(counter.data_ptr(), C.POINTER(C.c_uint)),
            C.cast(norms.data_ptr(), C.POINTER(C.c_double)), None, 0.0, s)
    assert rc == 0, fn
    return a_new, zc, norms


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_pass_equals_the_pass_with_caller_scratch(dtype):
    from spectral_galerkin_heat_amd import _lib
    lib = _lib.lib.load()
    code = _lib.F64 if dtype == torch.float64 else _lib.F32
    for (nx, ny, nz, nchunks, seed) in ((96, 64, 24, 4, 1), (40, 24, 8, 2, 2)):
        t = _case(nx, ny, nz, dtype, seed)
        n_pairs = lib.sgh_picard_partials_count(nx, ny, nz, nchunks)
        partials = torch.full((2 * n_pairs,), float("nan"), dtype=torch.float64, device="cuda")
        counter = torch.zeros(1, dtype=torch.int32, device="cuda")
        ref = _run(lib, "fused", t, nx, ny, nz, nchunks, code, dtype, 0.3, (partials, counter))
        got = _run(lib, "noalloc", t, nx, ny, nz, nchunks, code, dtype, 0.3)
        torch.cuda.synchronize()
        for a, b in zip(ref, got):
            assert torch.equal(a, b), (nx, ny, nz)
        assert torch.isfinite(got[2]).all() and float(got[2][0]) > 0.0
        assert int(counter.item()) == 0


def test_noalloc_refuses_more_workgroups_than_its_scratch():
    from spectral_galerkin_heat_amd import _lib
    lib = _lib.lib.load()
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = x.data_ptr()
    # 256 × 256 surface = 256 workgroups per chunk; 257 chunks exceed 65536 pairs
    rc = lib.sgh_picard_iter_fused_norms_noalloc(
        _lib.F64, 256, 256, 300, p, p, p, p, p, None, None, None, 0.1, None, 257, None, p,
        C.cast(p, C.POINTER(C.c_double)), None, 0.0, None)
    assert rc == -1 and "scratch" in _lib.last_error()

"""ctypes binding of libsgh_mi355x.so (the C ABI declared in include/sgh_mi355x.h).

The library is loaded from the package directory only; there is no fallback. Any
non-zero return code raises ``SGHError`` carrying ``sgh_last_error()``.
"""

from __future__ import annotations

import ctypes as C
import os

__all__ = ["SGHError", "lib", "call", "F64", "F32", "dtype_code", "ptr", "stream_ptr"]

F64, F32 = 0, 1
_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsgh_mi355x.so")

vp, i32, i64, f64, dp = C.c_void_p, C.c_int, C.c_longlong, C.c_double, C.POINTER(C.c_double)

# name -> argument types (all return int unless listed in _RESTYPES)
SIGNATURES = {
    "sgh_version": [],
    "sgh_last_error": [],
    "sgh_property_program_bytes": [],
    "sgh_etd_tables": [i32, i32, i32, i32, f64, f64, f64, f64, f64, f64, vp, vp, vp],
    "sgh_etd1_update": [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    "sgh_add_volume_source": [i32, i64, vp, vp, vp, vp],
    "sgh_add_surface_source": [i32, i32, i32, i32, vp, vp, vp, vp, vp],
    "sgh_mul_inplace": [i32, i64, vp, vp, vp],
    "sgh_zcontract": [i32, i32, i32, i32, vp, vp, i32, vp, vp],
    "sgh_picard_partials_count": [i32, i32, i32, i32],
    "sgh_picard_iter_fused": [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, f64, vp, i32,
                              vp, vp, vp, vp],
    "sgh_reduce_norms": [vp, i32, f64, vp, vp],
    "sgh_picard_iter_fused_norms": [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, f64, vp,
                                    i32, vp, vp, vp, vp, vp, vp, f64, vp],
    "sgh_picard_iter_fused_norms_noalloc": [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp,
                                            vp, f64, vp, i32, vp, vp, vp, vp, f64, vp],
    "sgh_host_alloc": [C.c_size_t, C.POINTER(vp)],
    "sgh_points_yz": [i32, i32, i32, i32, vp, vp, vp, vp, vp],
    "sgh_sep2d_forcing": [i32, i32, i32, vp, vp, vp, f64, vp, f64, f64, f64, f64, f64, f64, f64,
                          vp, f64, vp, vp, vp, vp],
    "sgh_host_free": [vp],
    "sgh_reduce": [i32, i64, vp, i32, vp, vp],
    "sgh_sep2d": [i32, i32, i32, i32, vp, vp, vp, i32, f64, vp, vp],
    "sgh_sum_parts": [i32, i64, i32, vp, vp],
    "sgh_axis_contract": [i32, i32, i64, i32, i64, vp, vp, f64, vp, vp],
    "sgh_axis_contract_rect": [i32, i32, i64, i32, i32, i64, vp, vp, f64, vp, vp],
    "sgh_axis_contract_strided": [i32, i32, i64, i32, i32, i64, vp, i64, vp, f64, vp, vp],
    "sgh_shift3d_xy": [i32, i32, i32, i32, vp, f64, f64, i32, vp, vp],
    "sgh_box_face_peaks": [i32, i32, i32, i32, vp, vp, vp],
    "sgh_plain_iter_size": [],
    "sgh_plain_iter_fork_init": [vp],
    "sgh_plain_iter_fork_release": [vp],
    "sgh_plain_iterate": [vp, i32, vp],
    "sgh_box_latent_source": [i32, i32, i32, i32, i32, vp, i32, i32, i32, vp, i64, vp, i64, vp,
                              vp, f64, f64, f64, f64, f64, vp, vp, vp, vp],
    "sgh_laser_flux": [i32, i32, i32, f64, f64, f64, f64, f64, f64, f64, f64, i32, vp, vp],
    "sgh_shift2d": [i32, i32, i32, vp, f64, f64, vp, f64, vp, vp],
    "sgh_surface_forcing": [i32, i64, vp, vp, f64, f64, f64, f64, f64, f64, f64, vp, f64, vp, vp,
                            vp],
    "sgh_convection_flux": [i32, i64, vp, f64, f64, vp, f64, vp, vp],
    "sgh_latent_source": [i32, i64, vp, vp, f64, f64, f64, f64, f64, vp, vp, vp],
    "sgh_property_correction_source": [i32, i32, i32, i32, vp, vp, vp, f64, f64, f64, f64, f64,
                                       f64, vp, vp, vp],
    "sgh_batch_pool_create": [i32, i32, i64, i64, i64, i32, i32, C.POINTER(vp)],
    "sgh_batch_create": [i32, i32, vp, i32, C.POINTER(vp), vp],
    "sgh_batch_add": [vp, vp, C.POINTER(i32), vp],
    "sgh_batch_remove": [vp, i32, vp],
    "sgh_batch_matrix_sizes": [vp, vp, i32],
    "sgh_batch_steps_done": [vp, vp, vp],
    "sgh_batch_destroy": [vp],
    "sgh_batch_run": [vp, i32, i32, vp],
    "sgh_batch_set_state": [vp, i32, vp, vp],
    "sgh_batch_get_state": [vp, i32, vp, vp],
    "sgh_batch_state_layout": [vp, i32, C.POINTER(i64), C.POINTER(i32), C.POINTER(i32),
                               C.POINTER(i32)],
    "sgh_batch_read_metrics": [vp, i32, vp, i32, vp],
    "sgh_batch_set_timing": [vp, i32],
    "sgh_batch_timing": [vp, vp],
    "sgh_batch_info": [vp, vp],
    "sgh_batch_variants": [vp, i32],
    "sgh_batch_set_variant": [vp, i32, i32, i32, vp],
    "sgh_batch_get_variant": [vp, vp],
    "sgh_batch_rewind": [vp, vp],
    "sgh_batch_set_loop": [vp, i32, vp],
    "sgh_batch_set_loop_timeout": [vp, f64, vp],
    "sgh_batch_loop_state": [vp, vp],
    "sgh_batch_set_team_max": [vp, i32, vp],
    "sgh_batch_team_sizes": [vp, vp, i32],
}


class CuboidSpec(C.Structure):
    """Mirror of sgh_cuboid_spec (include/sgh_mi355x.h)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("max_it", C.c_int),
        ("n_steps", C.c_int), ("cell_integrated", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("k", C.c_double), ("rho", C.c_double), ("cp", C.c_double),
        ("Pa", C.c_double), ("T_boil", C.c_double), ("dH", C.c_double), ("R_v", C.c_double),
        ("T_liq", C.c_double),
        ("absorptivity", C.c_double), ("rx", C.c_double), ("ry", C.c_double),
        ("order", C.c_double),
        ("dt", C.c_double), ("omega", C.c_double), ("tol", C.c_double),
        ("laser", C.POINTER(C.c_double)),
    ]


_RESTYPES = {"sgh_last_error": C.c_char_p, "sgh_plain_iter_size": C.c_size_t}


class SghPlainIter(C.Structure):
    """Mirror of ``SghPlainIter`` in include/sgh_mi355x.h (checked against
    ``sgh_plain_iter_size`` when the library loads)."""

    _fields_ = (
        [(n, C.c_int) for n in ("dtype", "xform", "nx", "ny", "nz", "nchunks", "n_pairs",
                                "latent_mode", "rescale", "have_bottom", "nxb", "nyb", "nzb",
                                "ntx", "nty", "x_start", "y_start", "pad0")]
        + [(n, C.c_double) for n in ("dxdy", "dV", "dV_fine", "n_elements", "omega", "Pa",
                                     "T_boil", "dH", "R_v", "T_liq", "h_top", "T0", "kbar",
                                     "abar", "T_S", "T_L", "rho", "L_f", "dt", "dx", "dy", "dz")]
        + [("aK", vp), ("KK", vp), ("a_buf", vp * 2), ("w_top", vp), ("w_bot", vp),
           ("zc", vp), ("T_s", vp), ("q_las", vp), ("q_evap", vp * 2), ("q_tmp", vp),
           ("S_top", vp), ("S_zero", vp), ("S_bot_raw", vp),
           ("sByT", vp), ("sBx", vp), ("sBy", vp), ("sBxT", vp),
           ("Q_latent", vp * 2), ("Q_modes", vp * 2), ("T_prev_full", vp), ("prog", vp),
           ("vBxT", vp), ("vByT", vp), ("vBzT", vp), ("vBx", vp), ("vBy", vp), ("vBz", vp),
           ("vol_w1", vp), ("vol_w2", vp), ("vol_T", vp),
           ("a_raw", vp), ("corr_work", vp), ("corr_f", vp), ("C_modes", vp),
           ("fBxT", vp), ("fByT", vp), ("fBzT", vp), ("fBx", vp), ("fBy", vp), ("fBz", vp),
           ("T_prev_box", vp), ("box_work", vp),
           ("partials", vp), ("norms", vp), ("norms_host", vp), ("counter", vp),
           ("seq", C.c_double), ("side_stream", vp), ("ev_fork", vp), ("ev_join", vp)])


class SGHError(RuntimeError):
    """A libsgh_mi355x call returned an error code."""


class DeviceOutOfMemory(SGHError):
    """A libsgh_mi355x call ran out of device memory (SGH_ERR_NOMEM from an allocation)."""


NOMEM = -4   # SGH_ERR_NOMEM


def is_device_oom(exc: BaseException) -> bool:
    """Whether ``exc`` (or an exception in its cause chain) is a device out-of-memory error:
    this library's DeviceOutOfMemory or torch's allocator error."""
    seen = set()
    while exc is not None and id(exc) not in seen:
        seen.add(id(exc))
        if isinstance(exc, DeviceOutOfMemory) or type(exc).__name__ == "OutOfMemoryError":
            return True
        exc = exc.__cause__ or exc.__context__
    return False


class _Lib:
    def __init__(self):
        self._h = None

    def load(self):
        if self._h is not None:
            return self._h
        # torch's HIP runtime must be the one in the process before this library binds
        # libamdhip64, so device buffers and streams are shared with torch.
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise SGHError(
                f"{LIB_PATH} is missing: build it with `python -m "
                "spectral_galerkin_heat_amd._build` (hipcc, gfx950). The MI355X backend has "
                "no CPU fallback.")
        h = C.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(h, name, None)
            if fn is None:
                continue
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, C.c_int)
        if hasattr(h, "sgh_plain_iter_size") and h.sgh_plain_iter_size() != C.sizeof(SghPlainIter):
            raise SGHError("SghPlainIter layout differs between the library "
                           f"({h.sgh_plain_iter_size()} bytes) and _lib.py "
                           f"({C.sizeof(SghPlainIter)} bytes): rebuild the library")
        self._h = h
        return h

    def __getattr__(self, name):
        return getattr(self.load(), name)


lib = _Lib()


def last_error() -> str:
    msg = lib.sgh_last_error()
    return msg.decode() if msg else ""


def call(name: str, *args) -> int:
    rc = getattr(lib, name)(*args)
    if rc < 0:
        msg = f"{name} failed ({rc}): {last_error()}"
        raise DeviceOutOfMemory(msg) if rc == NOMEM else SGHError(msg)
    return rc


def dtype_code(dtype) -> int:
    import numpy as np
    import torch
    if dtype in (torch.float64, np.float64) or str(dtype) in ("float64", "torch.float64"):
        return F64
    if dtype in (torch.float32, np.float32) or str(dtype) in ("float32", "torch.float32"):
        return F32
    raise ValueError(f"unsupported dtype {dtype!r} (float32 or float64)")


def ptr(t):
    """Device pointer of a tensor (None -> NULL). The tensor must be contiguous."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("MI355X kernels need C-contiguous tensors")
    return t.data_ptr()


def stream_ptr(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return s.cuda_stream

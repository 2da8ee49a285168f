"""The C-ABI library loads and exports every entry point declared in include/*.h.

Runs on the CPU: only symbol lookup and host-only entry points (no kernel launch).
"""

import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "spectral_galerkin_heat_amd", "libsgh_mi355x.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(sgh_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (python -m spectral_galerkin_heat_amd._build)")
    return ctypes.CDLL(LIB)


def test_every_declared_symbol_is_exported(lib):
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from spectral_galerkin_heat_amd import _lib
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_host_only_entry_points(lib):
    from spectral_galerkin_heat_amd.core.properties import PROGRAM_BYTES
    assert lib.sgh_version() >= 100
    assert lib.sgh_property_program_bytes() == PROGRAM_BYTES
    assert lib.sgh_picard_partials_count(64, 32, 16, 8) == 8 * 8
    assert lib.sgh_picard_partials_count(0, 32, 16, 8) < 0


def test_plain_iteration_descriptor_layout_matches_the_library():
    import ctypes
    from spectral_galerkin_heat_amd import _lib
    h = _lib.lib.load()
    assert h.sgh_plain_iter_size() == ctypes.sizeof(_lib.SghPlainIter)


@pytest.mark.skipif(os.path.exists("/dev/kfd"),
                    reason="argument checks only, run where no GPU could be launched on")
@pytest.mark.parametrize("outer,n,inner", [(65535, 8, (1 << 31) - 1), (4194240, (1 << 31) - 1, 1)])
def test_axis_contract_refuses_more_output_tiles_than_2_31(outer, n, inner):
    # refused on the host before any launch: the tile renumbering counts in 32 bits
    from spectral_galerkin_heat_amd import _lib
    h = _lib.lib.load()
    p = ctypes.c_void_p(256)
    rc = h.sgh_axis_contract_strided(_lib.F64, 0, outer, 8, n, inner, p, 8, ctypes.c_void_p(512),
                                     0.5, ctypes.c_void_p(1024), None)
    assert rc == -1
    assert b"output tiles" in h.sgh_last_error()


@pytest.mark.skipif(os.path.exists("/dev/kfd"),
                    reason="argument checks only, run where no GPU could be launched on")
def test_noalloc_pass_refuses_more_pairs_than_its_fixed_scratch():
    # refused on the host before the device is touched: 256×256 surface (256 workgroups per
    # z chunk) × 257 chunks exceeds the 65536 norm pairs of the library's fixed array
    from spectral_galerkin_heat_amd import _lib
    h = _lib.lib.load()
    p = ctypes.c_void_p(256)
    rc = h.sgh_picard_iter_fused_norms_noalloc(
        _lib.F64, 256, 256, 300, p, p, p, p, p, None, None, None, 0.1, None, 257, None,
        ctypes.c_void_p(512), ctypes.cast(1024, ctypes.POINTER(ctypes.c_double)), None, 0.0,
        None)
    assert rc == -1
    assert b"fixed scratch" in h.sgh_last_error()

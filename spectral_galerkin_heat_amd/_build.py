"""Build libsgh_mi355x.so in-tree with hipcc for gfx950.

Usage: ``python -m spectral_galerkin_heat_amd._build`` (or ``__graft_entry__.build()``).
Every HIP translation unit in ``csrc/`` is compiled for gfx950 with packed-FP32 ops
disabled and linked into one shared library next to this file.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libsgh_mi355x.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ARCH = os.environ.get("SGH_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]


def _hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the MI355X backend needs ROCm's hipcc to build")
    return path


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose: bool = False, jobs: int = 4, defines=(), libdir: str | None = None) -> str:
    """Compile every ``csrc/*.hip`` and link ``libsgh_mi355x.so``; return its path.

    ``defines`` (e.g. ``("SGH_RESIDENT_TIMESTAMPS",)``) and ``libdir`` build a diagnostic
    library into a directory of its own, leaving the product library alone."""
    hipcc = _hipcc()
    out_lib = LIB_PATH if libdir is None else os.path.join(libdir, LIB_NAME)
    objdir = os.path.join(HERE, "_build") if libdir is None else os.path.join(libdir, "_build")
    os.makedirs(objdir, exist_ok=True)
    dflags = [f"-D{d}" for d in defines]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [hipcc, *HIP_FLAGS, *dflags, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{r.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, sources()))
    tmp = out_lib + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out_lib)
    return out_lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

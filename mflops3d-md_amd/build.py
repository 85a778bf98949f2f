"""Build libmflops.so (HIP for gfx950) in-tree.

    python mflops3d-md_amd/build.py

Compiles every source under csrc/ with hipcc for gfx950 and links
lib/libmflops.so.  Always recompiles (a few tens of seconds).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# packed FP32 (v_pk_*_f32) stays off: the compiler only emits scalar FP32.
COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-result",
]
SOURCES = ["api.cpp", "setup.cpp", "io.cpp", "der.hip", "solve.hip", "ns.hip", "md.hip", "les.hip"]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    target = os.path.join(LIB, "libmflops.so")
    tmp = target + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, target)
    if verbose:
        print("built", target)
    build_fortran(verbose)
    return target


FC = os.environ.get("FC", "/opt/rocm/llvm/bin/amdflang")
FORTRAN = os.path.join(HERE, "fortran")


def build_fortran(verbose: bool = True):
    """ISO_C_BINDING modules + the Fortran driver (when amdflang is present)."""
    if not os.path.exists(FC):
        if verbose:
            print("amdflang not found: Fortran front end not built")
        return None
    out = os.path.join(FORTRAN, "_build")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(out, "test_solver_3d_mf")
    cmd = [FC, "-O2", "-J", out, os.path.join(FORTRAN, "mflops.f90"),
           os.path.join(FORTRAN, "test_solver_3d_mf.f90"), "-o", exe,
           "-L", LIB, "-lmflops", "-L", "/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../lib", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=out)
    if r.returncode != 0:
        raise RuntimeError("Fortran build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    if verbose:
        print("built", exe)
    return exe


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)

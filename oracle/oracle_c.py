"""ctypes front of the oracle's C restatements (oracle/residual.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
product path.

build() compiles residual.c with gcc (-O2, OpenMP, no fast-math: the 80-bit
sums must not be reassociated) into oracle/_build/liboracle_c.so; load()
returns the library, building it first when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "liboracle_c.so")
_lib = None


def build(verbose: bool = True) -> str:
    os.makedirs(OUT, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           os.path.join(HERE, "residual.c"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle C build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(verbose=False)
        lib = C.CDLL(LIB)
        lib.mr_residual.restype = C.c_int
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def multiplied_residual(sc, u_int, g_int, sigma, bcx=None, bcy=None, bcz=None) -> np.ndarray:
    """P (g - L u) of mf_oracle.multiplied_residual, evaluated in 80-bit
    arithmetic by residual.c and rounded to FP64; interior fields (n1,n2,n3)."""
    lib = load()
    u = np.asfortranarray(u_int, dtype=np.float64)
    g = np.asfortranarray(g_int, dtype=np.float64)
    n1, n2, n3 = u.shape
    if g.shape != u.shape:
        raise ValueError("u and g shapes differ")
    cs = (sc.cx, sc.cy, sc.cz)
    if tuple(c.n for c in cs) != (n1, n2, n3):
        raise ValueError("field shape does not match the solver coefficients")
    lhs = [np.ascontiguousarray(c.lhs, dtype=np.float64) for c in cs]
    rhs = [np.ascontiguousarray(c.rhs, dtype=np.float64) for c in cs]
    bw = np.ascontiguousarray(np.concatenate(
        [[c.ddr[0, 0], c.ddr[1, 0], c.ddr[c.n - 2, 7], c.ddr[c.n - 1, 6]] for c in cs]))
    shapes = ((n2, n3, 2), (n1, n3, 2), (n1, n2, 2))
    bcs = []
    for b, sh in zip((bcx, bcy, bcz), shapes):
        if b is None:
            bcs.append(None)
        else:
            b = np.asfortranarray(b, dtype=np.float64)
            if b.shape != sh:
                raise ValueError("boundary array shape %s != %s" % (b.shape, sh))
            bcs.append(b)
    res = np.zeros((n1, n2, n3), order="F")
    rc = lib.mr_residual(n1, n2, n3, *[_p(a) for a in lhs], *[_p(a) for a in rhs], _p(bw),
                         _p(u), _p(g), C.c_double(sigma),
                         *[(_p(b) if b is not None else None) for b in bcs], _p(res))
    if rc != 0:
        raise MemoryError("mr_residual: allocation failed")
    return res


def relative_residual(sc, u_full, g_full, sigma, bcx=None, bcy=None, bcz=None) -> float:
    """||P (g - L u)||_2 / ||P g_f||_2 on the interior of full-size fields, with
    P g_f = P (g - L 0) (the folded right-hand side multiplied through): the
    residual the mixed-precision solve reports, taken independently."""
    ui = np.asarray(u_full)[1:-1, 1:-1, 1:-1]
    gi = np.asarray(g_full)[1:-1, 1:-1, 1:-1]
    r = multiplied_residual(sc, ui, gi, sigma, bcx, bcy, bcz)
    b = multiplied_residual(sc, np.zeros_like(ui), gi, sigma, bcx, bcy, bcz)
    return float(np.linalg.norm(r.ravel()) / np.linalg.norm(b.ravel()))

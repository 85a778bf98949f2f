"""CPU tests of the C-ABI library: it loads, exports every symbol the header
declares, and its host-side setup (coefficient tables, eigen-decomposition)
agrees with the oracle.  No GPU calls."""
import os
import re

import numpy as np
import pytest

import mf_oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mflops", "mf.h")


def _lib():
    import mflops

    return mflops, mflops.load()


def test_header_symbols_exported():
    import ctypes

    mflops, lib = _lib()
    text = open(HEADER).read()
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mf_\w+)\s*\(", text, flags=re.M))
    assert len(names) >= 14, names
    for name in sorted(names):
        assert hasattr(lib, name), name
    from mflops._lib import SIGNATURES

    assert names == set(SIGNATURES), names ^ set(SIGNATURES)
    assert lib.mf_version() >= 100


def test_argument_errors_without_gpu():
    import ctypes as C

    mflops, lib = _lib()
    g = np.linspace(0, 1, 5)
    dl = np.zeros((5, 5), order="F")
    dr = np.zeros((5, 7), order="F")
    rc = lib.mf_der_coeffs_1d(g.ctypes.data_as(C.c_void_p), 5, 1, 6,
                              dl.ctypes.data_as(C.c_void_p), dr.ctypes.data_as(C.c_void_p))
    assert rc == 1 and b"invalid" in lib.mf_last_error()
    h = C.c_void_p()
    rc = lib.mf_solver3d_create(None, 4, None, 4, None, 4, None, 6, C.byref(h))
    assert rc == 1


@pytest.mark.parametrize("so", [4, 6, 8])
@pytest.mark.parametrize("der", [1, 2])
def test_der_coeffs_match_oracle(so, der):
    mflops, _ = _lib()
    g = o.grid_tanh(37, 0.0, 1.0, 1.7)
    dl, dr = mflops.der_coeffs_1d(g, der, so)
    odl, odr = o.der_coeffs_global(g, der, so)
    olu = o.pentalu(odl)
    np.testing.assert_allclose(dl, olu, rtol=1e-12, atol=1e-12 * np.abs(olu).max())
    np.testing.assert_allclose(dr, odr, rtol=1e-12, atol=1e-12 * np.abs(odr).max())


@pytest.mark.parametrize("bct", [(1, 1), (2, 2), (1, 2), (2, 1)])
@pytest.mark.parametrize("kind,n", [("u", 20), ("t", 64), ("a", 33)])
def test_solver_coeffs_match_oracle(bct, kind, n):
    mflops, _ = _lib()
    g = {"u": o.grid_uniform(n, 0, 1), "t": o.grid_tanh(n, 0, 1, 2.0),
         "a": o.grid_asin(n, -0.5, 0.5)}[kind]
    s = mflops.solver_coeffs_1d(g, bct, 6)
    c = o.SolverCoeffs1D(g, bct, 6)
    for key in ("ddl", "ddr", "neum"):
        ref = getattr(c, key)
        np.testing.assert_allclose(s[key], ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    np.testing.assert_allclose(s["inv_lhs"], c.inv_lhs, atol=1e-12 * np.abs(c.inv_lhs).max())
    # eigenvalues: same multiset as LAPACK dgeev (numpy), to the operator's scale
    scale = np.abs(c.eigen).max()
    np.testing.assert_allclose(np.sort(s["eigen"]), np.sort(c.eigen), atol=1e-11 * scale)
    # eigenpairs of the oracle's operator and an exact inverse
    A = c.operator()
    V = s["vect"]
    res = A @ V - V * s["eigen"][None, :]
    assert np.abs(res).max() < 1e-11 * scale
    np.testing.assert_allclose(s["inv_vect"] @ V, np.eye(n - 2), atol=1e-10)


def test_solver_complex_spectrum_is_an_error():
    """A strongly non-monotone grid can give a non-real spectrum; the library
    must refuse it like the reference (class_solver_3d.f90:537-540) or accept
    a real one -- never return garbage silently."""
    import ctypes as C

    mflops, lib = _lib()
    g = np.array([0, 0.01, 0.5, 0.51, 0.52, 0.9, 0.91, 1.3, 1.31, 2.0, 2.4, 2.41])
    try:
        s = mflops.solver_coeffs_1d(g, (1, 1), 6)
    except mflops.MFError as e:
        assert e.code == 2
        return
    c = o.SolverCoeffs1D(g, (1, 1), 6) if True else None
    A = c.operator()
    res = A @ s["vect"] - s["vect"] * s["eigen"][None, :]
    assert np.abs(res).max() < 1e-8 * np.abs(s["eigen"]).max()

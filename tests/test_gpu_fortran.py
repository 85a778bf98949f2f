"""The Fortran caller (ISO_C_BINDING modules over the C ABI) on the GPU.

Runs mflops3d-md_amd/fortran/_build/test_solver_3d_mf, which repeats the
reference's test_solver_3d (x^9+y^9+xy, sigma=2) and test_derivatives_omp
(alpha=1, beta=2) through the library, and checks each printed error
against the same quantity computed with the oracle; then the Fortran I/O
and NS modules: a field file round trip, the laminar channel as a steady
state of the NS step (to rounding) and a restart continued bitwise.
"""
import os
import subprocess

import numpy as np
import pytest

import mf_oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mflops3d-md_amd", "fortran", "_build", "test_solver_3d_mf")


def _oracle_errors(n):
    gx, gy, gz = (o.grid_asin(n, -0.5, 0.5), o.grid_asin(n, -0.5, 0.5), o.grid_asin(n, -0.25, 0.25))
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    uex = X ** 9 + Y ** 9 + X * Y
    g = np.asfortranarray(72 * X ** 7 + 72 * Y ** 7 + 2.0 * uex)
    bcx = np.asfortranarray(np.stack([uex[0, 1:-1, 1:-1], uex[-1, 1:-1, 1:-1]], -1))
    bcy = np.asfortranarray(np.stack([uex[1:-1, 0, 1:-1], uex[1:-1, -1, 1:-1]], -1))
    bcz = np.asfortranarray(np.stack([uex[1:-1, 1:-1, 0], uex[1:-1, 1:-1, -1]], -1))
    sc = o.SolverCoeffs3D(gx, gy, gz, (1,) * 6, 6)
    u = np.zeros_like(uex, order="F")
    o.solver_3d(sc, g, u, bcx, bcy, bcz, 2.0)
    d = u - uex
    s1 = (d[1:-1, 1:-1, 1:-1] ** 2).sum()
    s2 = (uex[1:-1, 1:-1, 1:-1] ** 2).sum()
    for sl in ((0, slice(1, -1), slice(1, -1)), (-1, slice(1, -1), slice(1, -1)),
               (slice(1, -1), 0, slice(1, -1)), (slice(1, -1), -1, slice(1, -1)),
               (slice(1, -1), slice(1, -1), 0), (slice(1, -1), slice(1, -1), -1)):
        s1 += (d[sl] ** 2).sum()
    out = {"solver_fp64_error": np.sqrt(s1) / np.sqrt(s2)}
    tf = lambda x: np.sin(2 * np.pi * x) + np.cos(4 * np.pi * x)  # noqa: E731
    f = np.asfortranarray(tf(X) + tf(Y) + tf(Z))
    for axis, c in enumerate((X, Y, Z)):
        dc = o.DerCoeffs((gx, gy, gz)[axis], solver=True, so=6)
        for order in (1, 2):
            if order == 1:
                ex = 2 * np.pi * (np.cos(2 * np.pi * c) - 2 * np.sin(4 * np.pi * c))
            else:
                ex = 4 * np.pi ** 2 * (-np.sin(2 * np.pi * c) - 4 * np.cos(4 * np.pi * c))
            df = o.derivative_3d(dc, f, axis, order)
            inner = (slice(1, -1),) * 3
            out["derivative_axis%d_order%d" % (axis, order)] = np.sqrt(
                ((df[inner] - ex[inner]) ** 2).sum() / (ex[inner] ** 2).sum())
    return out


@pytest.mark.gpu
def test_fortran_caller(tmp_path):
    assert os.path.exists(EXE), "Fortran driver not built (python mflops3d-md_amd/build.py)"
    n = 33
    print("run fortran driver n=%d" % n, flush=True)
    # the driver writes a field file and a restart directory into its cwd
    r = subprocess.run([EXE, str(n)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    got = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if len(parts) == 2:
            got[parts[0]] = float(parts[1])
    ref = _oracle_errors(n)
    # The same discretisation errors: two fields that agree to 1e-12 of their
    # max norm have relative L2 errors that differ by at most ~1e-12.
    for key, val in ref.items():
        assert key in got, key
        print("%s fortran %.16e oracle %.16e" % (key, got[key], val))
        assert abs(got[key] - val) <= 1e-11, (key, got[key], val)
    # field file round trip, and the NS module: a discrete steady state kept
    # to rounding, a restart continued bitwise
    assert got["io_roundtrip_maxdiff"] == 0.0 and got["io_dims_ok"] == 1
    assert got["ns_channel_steady"] <= 1e-11
    assert abs(got["ns_time"] - 5e-3) <= 1e-15
    assert got["ns_restart_maxdiff"] == 0.0
    assert sorted(os.listdir(tmp_path / "restart_f"))[0] == "param.dat"
    # der/dder and der_s/dder_s line by line, as the reference loops over a
    # field's lines, agree with the 3-D driver (same schemes; the line kernel
    # solves each line sequentially, the 3-D driver in segments, so second
    # derivatives differ by rounding, measured 6.5e-13 of the largest value;
    # one launch per line through the handle's staging buffer, no
    # allocation per call)
    for order in (1, 2):
        assert got["der_lines_order%d_vs_3d" % order] <= 1e-12
        print("der line loop order %d: %.1f us per line" % (order, got["der_lines_order%d_us_per_line" % order]))
    assert got["solver_mixed_vs_fp64"] <= 1e-10
    assert got["solver_mixed_residual"] <= 1e-12
    assert abs(got["solver_mixed_error"] - ref["solver_fp64_error"]) <= 1e-11

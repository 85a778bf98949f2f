"""The reference's own FP64 solver and compact derivatives, run as programs
(test and baseline infrastructure only, never the product path).

`oracle/build_ref.sh` compiles the reference's solver sources
(src/class_solver/class_solver_3d.f90, class_solver_coefficient_3d.f90,
src/precision/precision.f90) with the driver `oracle/ref_solver_driver.f90`
into `oracle/_ref/ref_solver_driver`, and its derivative sources
(src/class_derivatives/class_derivatives.f90,
class_derivatives_coefficient.f90) with `oracle/ref_der_driver.f90` into
`oracle/_ref/ref_der_driver`.  `solve` writes a problem, runs the solver and
reads back the solution of solve_3d (class_solver_3d.f90:83) and its time per
call; `derivatives` returns der, dder, der_s and dder_s
(class_derivatives.f90:186-260) of a set of lines.
"""
from __future__ import annotations

import os
import resource
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "_ref", "ref_solver_driver")
DER_EXE = os.path.join(HERE, "_ref", "ref_der_driver")
BUILD = os.path.join(HERE, "build_ref.sh")
REF = "/root/reference"


def _big_stack():
    # solve_3d keeps three field-sized work arrays as automatic (stack) arrays
    hard = resource.getrlimit(resource.RLIMIT_STACK)[1]
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def available() -> bool:
    return os.access(EXE, os.X_OK)


def der_available() -> bool:
    return os.access(DER_EXE, os.X_OK)


def build() -> bool:
    """Build the driver when the reference's sources are present; returns
    whether it is available afterwards."""
    if os.path.isdir(REF):
        subprocess.run(["sh", BUILD], check=True, capture_output=True)
    return available()


def solve(gx, gy, gz, g, sigma: float, bcx=None, bcy=None, bcz=None, bct=(1,) * 6,
          so: int = 6, nrep: int = 1, threads: int | None = None):
    """solve_3d of the reference on g (full size, Fortran order; its interior
    is the right-hand side) with face values bcx (ny-2, nz-2, 2), bcy, bcz
    (zero when None).  Returns (u full size, seconds per solve_3d call,
    seconds of solver_coeffs_init_3d)."""
    if not available():
        raise RuntimeError("reference solver not built: run `sh oracle/build_ref.sh`")
    nx, ny, nz = (len(gx), len(gy), len(gz))
    if g.shape != (nx, ny, nz):
        raise ValueError("g must be (%d, %d, %d)" % (nx, ny, nz))
    n1, n2, n3 = nx - 2, ny - 2, nz - 2
    faces = []
    for f, shape in ((bcx, (n2, n3, 2)), (bcy, (n1, n3, 2)), (bcz, (n1, n2, 2))):
        faces.append(np.zeros(shape) if f is None else np.asarray(f, dtype=np.float64))
        if faces[-1].shape != shape:
            raise ValueError("face array of shape %s, expected %s" % (faces[-1].shape, shape))
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = env["MKL_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            np.array([nx, ny, nz, so, nrep] + list(bct), dtype=np.int32).tofile(f)
            np.array([sigma], dtype=np.float64).tofile(f)
            for a in (gx, gy, gz):
                np.asarray(a, dtype=np.float64).tofile(f)
            np.asarray(g[1:-1, 1:-1, 1:-1], dtype=np.float64).ravel(order="F").tofile(f)
            for a in faces:
                a.ravel(order="F").tofile(f)
        subprocess.run([EXE, fin, fout], check=True, env=env, preexec_fn=_big_stack)
        out = np.fromfile(fout, dtype=np.float64)
    t_solve, t_setup = float(out[0]), float(out[1])
    u = np.asfortranarray(out[2:].reshape((nx, ny, nz), order="F"))
    return u, t_solve, t_setup


def derivatives(grid, x, so: int = 6, solver: bool = True):
    """The reference's der, dder (and, in solver mode, der_s, dder_s) of the
    lines x (n, nlines), grid of n points.  Returns a dict of (n, nlines)
    arrays; der_s and dder_s hold 0 at the first and last points."""
    if not der_available():
        raise RuntimeError("reference derivatives not built: run `sh oracle/build_ref.sh`")
    grid = np.asarray(grid, dtype=np.float64)
    n = len(grid)
    x = np.asarray(x, dtype=np.float64).reshape(n, -1, order="F")
    nl = x.shape[1]
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            np.array([n, so, 1 if solver else 0, nl], dtype=np.int32).tofile(f)
            grid.tofile(f)
            x.ravel(order="F").tofile(f)
        subprocess.run([DER_EXE, fin, fout], check=True, preexec_fn=_big_stack)
        out = np.fromfile(fout, dtype=np.float64)
    names = ["der", "dder"] + (["der_s", "dder_s"] if solver else [])
    return {k: out[i * n * nl:(i + 1) * n * nl].reshape((n, nl), order="F")
            for i, k in enumerate(names)}

"""The numpy oracle's solver against the reference's own solve_3d
(class_solver_3d.f90:83-297, with solver_coeffs_init_3d :380-466), compiled
from the reference's sources by oracle/build_ref.sh (MKL's LAPACK/BLAS) into
oracle/_ref/.  Dirichlet and mixed Dirichlet/Neumann boundaries with nonzero
face values, a stretched grid, Helmholtz and Poisson sigma.  Skipped when the
reference build is absent."""
import numpy as np
import pytest

import mf_oracle as o
import ref_solver

CASES = [
    ((24, 20, 16), (1,) * 6, -300.0),
    ((22, 26, 18), (1, 2, 1, 1, 2, 2), -50.0),
    ((30, 28, 24), (1,) * 6, 0.0),
    ((18, 34, 20), (2, 1, 1, 2, 1, 1), -4.5e3),
]


@pytest.mark.skipif(not ref_solver.available(),
                    reason="reference build absent (oracle/_ref: sh oracle/build_ref.sh)")
@pytest.mark.parametrize("shape,bct,sigma", CASES)
def test_oracle_solver_matches_reference_build(shape, bct, sigma):
    gx = o.grid_uniform(shape[0], 0.0, 1.0)
    gy = o.grid_tanh(shape[1], 0.0, 1.0, 1.5)
    gz = o.grid_uniform(shape[2], 0.0, 2.0)
    rng = np.random.default_rng(7)
    g = np.asfortranarray(rng.standard_normal(shape))
    n1, n2, n3 = (s - 2 for s in shape)
    faces = [np.asfortranarray(rng.standard_normal(f)) for f in ((n2, n3, 2), (n1, n3, 2), (n1, n2, 2))]
    sc = o.SolverCoeffs3D(gx, gy, gz, bct, 6)
    u = np.zeros(shape, order="F")
    o.solver_3d(sc, g, u, *faces, sigma)
    ref, _, _ = ref_solver.solve(gx, gy, gz, g, sigma, *faces, bct=bct, threads=2)
    scale = np.abs(ref).max()
    inner = (slice(1, -1),) * 3
    # interior: the solve itself; the faces: the Neumann extrapolation too
    assert np.abs(u[inner] - ref[inner]).max() <= 1e-13 * scale
    assert np.abs(u - ref).max() <= 1e-12 * scale

"""The numpy oracle against the reference's own code, compiled from the
reference's sources by oracle/build_ref.sh (MKL's LAPACK/BLAS) into
oracle/_ref/:

* solve_3d (class_solver_3d.f90:83-297, with solver_coeffs_init_3d
  :380-466): Dirichlet and mixed Dirichlet/Neumann boundaries with nonzero
  face values, a stretched grid, Helmholtz and Poisson sigma;
* der, dder, der_s, dder_s (class_derivatives.f90:111-260) on uniform and
  tanh-stretched grids, orders so = 4, 6, 8.

Skipped when the reference build is absent."""
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


DER_CASES = [(64, "uniform", 6), (40, "tanh", 6), (33, "tanh", 4), (192, "tanh", 6),
             (12, "tanh", 6), (50, "uniform", 8)]


@pytest.mark.skipif(not ref_solver.der_available(),
                    reason="reference build absent (oracle/_ref: sh oracle/build_ref.sh)")
@pytest.mark.parametrize("n,grid,so", DER_CASES)
def test_oracle_derivatives_match_reference_build(n, grid, so):
    """Both sides solve the same Taylor systems for the scheme coefficients
    with LAPACK in FP64 (numpy's OpenBLAS, MKL); at so = 8 those systems are
    ill-conditioned enough that the two LAPACKs' coefficients differ near
    1e-12 relative (the oracle's own are pinned against 40-digit arithmetic
    in test_oracle.py), hence the wider bar there."""
    g = o.grid_uniform(n, 0.0, 1.0) if grid == "uniform" else o.grid_tanh(n, 0.0, 1.0, 2.0)
    x = np.random.default_rng(n).standard_normal((n, 9))
    ref = ref_solver.derivatives(g, x, so=so, solver=True)
    dc = o.DerCoeffs(g, True, so)
    bar = 1e-12 if so <= 6 else 2e-11
    for name in ("der", "dder", "der_s", "dder_s"):
        rows = slice(1, -1) if name.endswith("_s") else slice(None)
        got, want = getattr(dc, name)(x)[rows], ref[name][rows]
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= bar, (name, err)

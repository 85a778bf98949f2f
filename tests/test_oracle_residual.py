"""Pins of the residual yardstick behind the mixed-precision parity checks.

mf_oracle.multiplied_residual (numpy, 80-bit) and its C restatement
oracle/residual.c (80-bit, OpenMP; the full-size checker) both evaluate
P (g - L u), P = lhs_x lhs_y lhs_z, of the operator class_solver_3d.f90:104-225
inverts.  They are pinned here by:
* a dense route: P and L assembled as Kronecker products of the dense 1-D
  matrices (L_a = inv(lhs_a) rhs_a, boundary columns from a unit boundary
  value through the dense operator) on an 11x12x10 grid with mixed boundary
  codes, stretched grids (non-symmetric lhs/rhs: a transposed factor fails),
  random u, g and boundary arrays (a dropped or mis-signed boundary term
  fails);
* the C and numpy evaluations agreeing to rounding on several grids;
* the residual of a solution refined to ~1e-15 being ~1e-15, and of a
  perturbed one growing with the perturbation.
"""
import numpy as np
import pytest

import mf_oracle as o
import oracle_c as oc


def _grids(shape):
    return (o.grid_uniform(shape[0], 0.0, 1.0), o.grid_tanh(shape[1], 0.0, 1.0, 1.5),
            o.grid_asin(shape[2], 0.0, 0.5))


def _random(shape, seed):
    rng = np.random.default_rng(seed)
    n1, n2, n3 = (s - 2 for s in shape)
    u = rng.standard_normal((n1, n2, n3))
    g = rng.standard_normal((n1, n2, n3))
    bcs = (rng.standard_normal((n2, n3, 2)), rng.standard_normal((n1, n3, 2)),
           rng.standard_normal((n1, n2, 2)))
    return u, g, [np.asfortranarray(b) for b in bcs]


def _dense_residual(sc, u, g, sigma, bcs):
    """P (g - L u) with every operator a dense Kronecker product (float64)."""
    cs = (sc.cx, sc.cy, sc.cz)
    A = [c.inv_lhs @ c.rhs for c in cs]
    I = [np.eye(c.n) for c in cs]
    K = (np.kron(I[2], np.kron(I[1], A[0])) + np.kron(I[2], np.kron(A[1], I[0]))
         + np.kron(A[2], np.kron(I[1], I[0])) + sigma * np.eye(u.size))
    P = np.kron(cs[2].lhs, np.kron(cs[1].lhs, cs[0].lhs))
    # boundary part: the boundary columns of each 1-D operator, written out
    # from the solver's ddr weights on the two rows next to each face
    bpart = np.zeros(u.shape)
    for a, c in enumerate(cs):
        n = c.n
        col1 = np.zeros(n)
        col2 = np.zeros(n)
        col1[0], col1[1] = c.ddr[0, 0], c.ddr[1, 0]
        col2[n - 2], col2[n - 1] = c.ddr[n - 2, 7], c.ddr[n - 1, 6]
        e1, e2 = c.inv_lhs @ col1, c.inv_lhs @ col2
        b = bcs[a]
        t = e1[:, None, None] * b[None, :, :, 0] + e2[:, None, None] * b[None, :, :, 1]
        bpart += np.moveaxis(t, 0, a)
    r = g.ravel(order="F") - K @ u.ravel(order="F") - bpart.ravel(order="F")
    return (P @ r).reshape(u.shape, order="F")


@pytest.mark.parametrize("bct,sigma", [((1,) * 6, 2.0), ((1, 2, 2, 1, 1, 2), -300.0),
                                       ((2,) * 6, 0.0)])
def test_multiplied_residual_vs_dense(bct, sigma):
    shape = (11, 12, 10)
    sc = o.SolverCoeffs3D(*_grids(shape), bct, 6)
    u, g, bcs = _random(shape, 3)
    ref = _dense_residual(sc, u, g, sigma, bcs)
    num = np.asarray(o.multiplied_residual(sc, u, g, sigma, *bcs), dtype=np.float64)
    cres = oc.multiplied_residual(sc, u, g, sigma, *bcs)
    scale = np.abs(ref).max()
    e_num = np.abs(num - ref).max() / scale
    e_c = np.abs(cres - ref).max() / scale
    print("numpy vs dense %.2e, C vs dense %.2e" % (e_num, e_c))
    assert e_num < 1e-12 and e_c < 1e-12
    # a mistake planted in the dense route (one transposed lhs) is caught
    cs = (sc.cx, sc.cy, sc.cz)
    lhs_y = cs[1].lhs.copy()
    try:
        cs[1].lhs = lhs_y.T.copy()
        bad = _dense_residual(sc, u, g, sigma, bcs)
    finally:
        cs[1].lhs = lhs_y
    assert np.abs(bad - ref).max() / scale > 1e-3


@pytest.mark.parametrize("shape,bct,sigma", [((9, 10, 11), (1,) * 6, 50.0),
                                             ((20, 14, 17), (2, 1, 1, 2, 2, 1), -1e3),
                                             ((33, 31, 12), (2,) * 6, 0.0)])
def test_c_residual_matches_numpy(shape, bct, sigma):
    sc = o.SolverCoeffs3D(*_grids(shape), bct, 6)
    u, g, bcs = _random(shape, 11)
    num = np.asarray(o.multiplied_residual(sc, u, g, sigma, *bcs), dtype=np.float64)
    cres = oc.multiplied_residual(sc, u, g, sigma, *bcs)
    err = np.abs(cres - num).max() / np.abs(num).max()
    print("C vs numpy (80-bit both) %.2e" % err)
    assert err < 1e-15


def test_relative_residual_of_refined_solution():
    shape = (18, 22, 16)
    sc = o.SolverCoeffs3D(*_grids(shape), (1,) * 6, 6)
    rng = np.random.default_rng(5)
    g = np.asfortranarray(rng.standard_normal(shape))
    bcs = [np.asfortranarray(rng.standard_normal(s)) for s in ((20, 14, 2), (16, 14, 2), (16, 20, 2))]
    sigma = -40.0
    ui = o.solve_exact(sc, g, *bcs, sigma)
    u = np.zeros(shape, order="F")
    u[1:-1, 1:-1, 1:-1] = np.asarray(ui, dtype=np.float64)
    r0 = oc.relative_residual(sc, u, g, sigma, *bcs)
    # the same field perturbed by 1e-10 relative leaves a residual of that order
    up = u.copy()
    up[1:-1, 1:-1, 1:-1] *= 1.0 + 1e-10 * rng.standard_normal((16, 20, 14))
    r1 = oc.relative_residual(sc, up, g, sigma, *bcs)
    print("refined %.2e, perturbed %.2e" % (r0, r1))
    assert r0 < 1e-15
    assert 1e-12 < r1 < 1e-8

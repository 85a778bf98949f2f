"""GPU parity: the HIP path through the C ABI against the oracle.

Bar: FP64 derivatives and solves within 1e-12 of the oracle relative to the
field's max norm; the mixed-precision solve reaches a relative residual
<= 1e-12 of the FP64 operator and lands within 1e-12 of the oracle.
"""
import numpy as np
import pytest

import mf_oracle as o

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _mf():
    import mflops
    import torch

    assert torch.cuda.is_available(), "GPU test run without a GPU"
    mflops.load()
    return mflops, torch


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ------------------------------------------------------------- config 1 ----

def c1_problem():
    """64^3 box [0,1]^3, tanh stretch 1.5 in y, uniform x,z, Dirichlet;
    u = sin(pi x) sin(pi y) sin(pi z) + 1e-3 N(0,1) (seed 0)."""
    n = 64
    gx = o.grid_uniform(n, 0.0, 1.0)
    gy = o.grid_tanh(n, 0.0, 1.0, 1.5)
    gz = o.grid_uniform(n, 0.0, 1.0)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    noise = o.normal(0, n ** 3).reshape((n, n, n), order="F")
    u = np.asfortranarray(np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z) + 1e-3 * noise)
    return gx, gy, gz, u


@pytest.mark.parametrize("solver_mode", [False, True])
def test_derivatives_c1(solver_mode):
    mf, torch = _mf()
    gx, gy, gz, u = c1_problem()
    ud = mf.to_device(u)
    for axis, g in enumerate((gx, gy, gz)):
        dc = mf.DerivativeCoefficients(g, solver=solver_mode, so=6)
        odc = o.DerCoeffs(g, solver=solver_mode, so=6)
        for order in (1, 2):
            print("derivative c1 axis=%d order=%d solver=%s" % (axis, order, solver_mode), flush=True)
            got = mf.to_host(mf.derivative(dc, ud, axis, order))
            torch.cuda.synchronize()
            ref = o.derivative_3d(odc, u, axis, order)
            assert _rel(got, ref) <= TOL, (axis, order, _rel(got, ref))


@pytest.mark.parametrize("shape", [(9, 10, 11), (33, 17, 40), (70, 65, 8 + 2)])
@pytest.mark.parametrize("so", [4, 6, 8])
def test_derivatives_shapes(shape, so):
    mf, torch = _mf()
    rng = np.random.default_rng(sum(shape) + so)
    gs = [o.grid_uniform(shape[0], 0, 1), o.grid_tanh(shape[1], 0, 2, 1.8),
          o.grid_asin(shape[2], -0.5, 0.5)]
    # smooth field + 1e-3 noise, as the reference's derivative tests
    # (test_derivatives_omp.f90:65-99) and the 64^3 configuration
    X, Y, Z = np.meshgrid(*gs, indexing="ij")
    f = np.asfortranarray(np.sin(2 * np.pi * X) + np.cos(np.pi * Y) * np.sin(4 * Z)
                          + 1e-3 * rng.standard_normal(shape))
    fd = mf.to_device(f)
    for axis in range(3):
        for solver in (False, True):
            if solver and shape[axis] < 10:
                continue
            dc = mf.DerivativeCoefficients(gs[axis], solver=solver, so=so)
            odc = o.DerCoeffs(gs[axis], solver=solver, so=so)
            for order in (1, 2):
                print("derivative shape=%s so=%d axis=%d order=%d solver=%s"
                      % (shape, so, axis, order, solver), flush=True)
                got = mf.to_host(mf.derivative(dc, fd, axis, order))
                ref = o.derivative_3d(odc, f, axis, order)
                assert _rel(got, ref) <= TOL, (axis, order, solver, _rel(got, ref))


def test_der_lines():
    mf, torch = _mf()
    g = o.grid_tanh(50, 0, 1, 1.3)
    dc = mf.DerivativeCoefficients(g, solver=True, so=6)
    odc = o.DerCoeffs(g, solver=True, so=6)
    rng = np.random.default_rng(5)
    x = np.asfortranarray(rng.standard_normal((50, 37)))
    xd = mf.to_device(x)
    for name in ("der", "dder", "der_s", "dder_s"):
        print("der_lines", name, flush=True)
        out = mf.fortran_zeros((50, 37))
        getattr(dc, name)(xd, out)
        got = mf.to_host(out)
        ref = getattr(odc, name)(x)
        assert _rel(got, ref) <= TOL, name


# ---------------------------------------------------------------- solver ---

def _make_problem(shape, bct, sigma, seed, grids=None):
    """Random interior u and bc; g = operator(u) so the discrete answer is u."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    if grids is None:
        grids = (o.grid_uniform(nx, 0, 1), o.grid_tanh(ny, 0, 1, 1.5), o.grid_asin(nz, 0, 0.5))
    sc = o.SolverCoeffs3D(*grids, bct, 6)
    uin = rng.standard_normal((nx - 2, ny - 2, nz - 2))
    bcx = np.asfortranarray(rng.standard_normal((ny - 2, nz - 2, 2)))
    bcy = np.asfortranarray(rng.standard_normal((nx - 2, nz - 2, 2)))
    bcz = np.asfortranarray(rng.standard_normal((nx - 2, ny - 2, 2)))
    g = np.zeros(shape, order="F")
    g[1:-1, 1:-1, 1:-1] = o.apply_operator(sc, uin, sigma, bcx, bcy, bcz)
    return grids, sc, g, (bcx, bcy, bcz), uin


def _solve_gpu(mf, grids, bct, g, bcs, sigma, precision, u0=None):
    sc = mf.SolverCoeffs3D(*grids, bct, 6)
    gd = mf.to_device(g)
    u0 = np.full(g.shape, 7.0, order="F") if u0 is None else u0
    ud = mf.to_device(u0)
    bc = mf.BoundaryCondition(*[mf.to_device(b) for b in bcs])
    rep = sc.solve(gd, ud, bc, sigma, precision)
    return mf.to_host(ud), rep


CASES = [
    ((11, 12, 10), (1,) * 6, 2.0),
    ((11, 12, 10), (2,) * 6, 5.0),
    ((11, 12, 10), (1, 2, 2, 1, 1, 2), -40.0),
    ((34, 29, 40), (1,) * 6, 1e3),
    ((34, 29, 40), (1, 1, 2, 2, 1, 1), -300.0),
    ((66, 70, 21), (1,) * 6, 0.0),
]


@pytest.mark.parametrize("shape,bct,sigma", CASES)
def test_solve_fp64_vs_oracle(shape, bct, sigma):
    mf, torch = _mf()
    grids, sc, g, bcs, uin = _make_problem(shape, bct, sigma, 11)
    print("solve fp64 shape=%s bct=%s sigma=%g" % (shape, bct, sigma), flush=True)
    got, _ = _solve_gpu(mf, grids, bct, g, bcs, sigma, mf.MF_FP64)
    ref = np.full(shape, 7.0, order="F")
    o.solver_3d(sc, g, ref, *bcs, sigma)
    assert _rel(got, ref) <= TOL, _rel(got, ref)
    assert np.all(got[0, 0, :] == 7.0) and np.all(got[:, -1, -1] == 7.0)  # edges untouched


@pytest.mark.parametrize("shape,bct,sigma", CASES)
def test_solve_mixed_vs_oracle(shape, bct, sigma):
    mf, torch = _mf()
    grids, sc, g, bcs, uin = _make_problem(shape, bct, sigma, 12)
    print("solve mixed shape=%s bct=%s sigma=%g" % (shape, bct, sigma), flush=True)
    got, rep = _solve_gpu(mf, grids, bct, g, bcs, sigma, mf.MF_MIXED)
    assert rep.residual <= 1e-12 and rep.iterations >= 2, (rep.iterations, rep.residual)
    ref = np.full(shape, 7.0, order="F")
    o.solver_3d(sc, g, ref, *bcs, sigma)
    assert _rel(got, ref) <= TOL, _rel(got, ref)
    # true residual with the oracle's (unmultiplied) operator
    res = o.apply_operator(sc, got[1:-1, 1:-1, 1:-1], sigma, *bcs) - g[1:-1, 1:-1, 1:-1]
    assert np.abs(res).max() <= 1e-10 * np.abs(g).max()


def test_solve_c1_helmholtz_and_poisson():
    """Config 1: sigma=1e3 Helmholtz and sigma=0 Poisson, RHS = operator(u)."""
    mf, torch = _mf()
    gx, gy, gz, u = c1_problem()
    bct = (1,) * 6
    sc = o.SolverCoeffs3D(gx, gy, gz, bct, 6)
    bcx = np.asfortranarray(np.stack([u[0, 1:-1, 1:-1], u[-1, 1:-1, 1:-1]], -1))
    bcy = np.asfortranarray(np.stack([u[1:-1, 0, 1:-1], u[1:-1, -1, 1:-1]], -1))
    bcz = np.asfortranarray(np.stack([u[1:-1, 1:-1, 0], u[1:-1, 1:-1, -1]], -1))
    for sigma in (1e3, 0.0):
        g = np.zeros_like(u)
        g[1:-1, 1:-1, 1:-1] = o.apply_operator(sc, u[1:-1, 1:-1, 1:-1], sigma, bcx, bcy, bcz)
        ref = np.zeros_like(u)
        o.solver_3d(sc, g, ref, bcx, bcy, bcz, sigma)
        for prec in (mf.MF_FP64, mf.MF_MIXED):
            print("c1 solve sigma=%g precision=%d" % (sigma, prec), flush=True)
            got, rep = _solve_gpu(mf, (gx, gy, gz), bct, g, (bcx, bcy, bcz), sigma, prec,
                                  u0=np.zeros_like(u))
            m = np.ones(u.shape, bool)
            assert _rel(got, ref) <= TOL, (sigma, prec, _rel(got, ref))
            inner = (slice(1, -1),) * 3
            assert _rel(got[inner], u[inner]) <= 1e-11

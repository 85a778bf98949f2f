"""GPU parity: the HIP path through the C ABI against the oracle.

Two FP64 implementations of the same algorithm differ by their rounding; for
the strongly cancelling second derivatives and for the diagonalisation solve
the FP64 oracle's own error reaches 1e-12 .. 1e-11 (measured below).  Parity is
therefore measured against the oracle evaluated exactly (80-bit arithmetic for
the derivative algorithm; the discrete solution refined to ~1e-15 for solves):

* derivatives and FP64 solves: |gpu - exact| <= max(1e-12, 2 |oracle64 - exact|)
  relative to max|exact|, i.e. 1e-12 or no worse than the FP64 oracle;
* mixed-precision solves: relative residual of the FP64 operator <= tol; at
  tol 1e-12 |gpu - exact| <= 2e-11, at tol 1e-14 |gpu - exact| <= max(1e-12,
  |oracle64 - exact|): FP64 refinement stalls near cond * eps64, so the mixed
  solve is held to 1e-12 or to the FP64 reference algorithm's own error.
Every test prints the three distances.
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


def _check(label, got, exact, ora, tol=TOL, allow_oracle=True, ora_factor=2.0):
    ex = np.asarray(exact, dtype=np.float64)
    sc = max(np.abs(ex).max(), 1e-300)
    e_gpu = np.abs(got - ex).max() / sc
    e_ora = np.abs(ora - ex).max() / sc
    e_go = np.abs(got - ora).max() / sc
    bar = max(tol, ora_factor * e_ora) if allow_oracle else tol
    print("%s: |gpu-exact| %.2e  |oracle64-exact| %.2e  |gpu-oracle64| %.2e  (bar %.1e)"
          % (label, e_gpu, e_ora, e_go, bar), flush=True)
    assert e_gpu <= bar, (label, e_gpu, e_ora)


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


def _der_case(mf, grids, f, so, solvers=(False, True), label=""):
    fd = mf.to_device(f)
    for axis in range(3):
        for solver in solvers:
            if solver and f.shape[axis] < 10:
                continue
            dc = mf.DerivativeCoefficients(grids[axis], solver=solver, so=so)
            odc = o.DerCoeffs(grids[axis], solver=solver, so=so)
            for order in (1, 2):
                tag = "%s so=%d axis=%d order=%d solver=%s" % (label, so, axis, order, solver)
                print("run " + tag, flush=True)
                got = mf.to_host(mf.derivative(dc, fd, axis, order))
                ora = o.derivative_3d(odc, f, axis, order)
                ex = o.derivative_3d_extended(odc, f, axis, order)
                _check(tag, got, ex, ora)


def test_derivatives_c1():
    mf, torch = _mf()
    gx, gy, gz, u = c1_problem()
    _der_case(mf, (gx, gy, gz), u, 6, label="c1")


# line lengths cover one to four 32-row segments (9 .. 130, a 38-row last
# segment at 70, sub-lines with fewer segments than the full lines at 39) and
# a 340-point line, longer than the segmented solve's LDS tables allow, which
# takes the one-lane-per-line kernel
@pytest.mark.parametrize("shape", [(9, 10, 11), (33, 17, 40), (70, 65, 10), (130, 12, 9),
                                   (39, 340, 38)])
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
    _der_case(mf, gs, f, so, label=str(shape))


def test_derivatives_white_noise():
    """Pure noise input: the worst case for cancellation."""
    mf, torch = _mf()
    shape = (40, 33, 12)
    rng = np.random.default_rng(3)
    gs = [o.grid_tanh(shape[0], 0, 1, 2.0), o.grid_uniform(shape[1], 0, 3),
          o.grid_asin(shape[2], 0, 1)]
    f = np.asfortranarray(rng.standard_normal(shape))
    _der_case(mf, gs, f, 6, label="noise")


def test_der_lines():
    mf, torch = _mf()
    g = o.grid_tanh(50, 0, 1, 1.3)
    dc = mf.DerivativeCoefficients(g, solver=True, so=6)
    odc = o.DerCoeffs(g, solver=True, so=6)
    rng = np.random.default_rng(5)
    x = np.asfortranarray(rng.standard_normal((50, 37)))
    xd = mf.to_device(x)
    for name in ("der", "dder", "der_s", "dder_s"):
        out = mf.fortran_zeros((50, 37))
        getattr(dc, name)(xd, out)
        got = mf.to_host(out)
        ora = getattr(odc, name)(x)
        ex = getattr(odc, name)(x.astype(np.longdouble))
        _check("lines " + name, got, ex, ora)


@pytest.mark.parametrize("n,so", [(64, 6), (192, 6), (33, 4)])
def test_der_lines_match_reference_build(n, so):
    """der, dder, der_s, dder_s on the GPU against the reference's own
    (class_derivatives.f90:186-260, compiled from its sources by
    oracle/build_ref.sh) on a tanh-stretched grid."""
    import ref_solver
    if not ref_solver.der_available():
        pytest.skip("reference build absent (oracle/_ref: sh oracle/build_ref.sh)")
    mf, torch = _mf()
    print("run der lines vs reference build n=%d so=%d" % (n, so), flush=True)
    g = o.grid_tanh(n, 0, 1, 2.0)
    dc = mf.DerivativeCoefficients(g, solver=True, so=so)
    x = np.asfortranarray(np.random.default_rng(n).standard_normal((n, 45)))
    ref = ref_solver.derivatives(g, x, so=so, solver=True)
    xd = mf.to_device(x)
    for name in ("der", "dder", "der_s", "dder_s"):
        out = mf.fortran_zeros((n, 45))
        getattr(dc, name)(xd, out)
        rows = slice(1, -1) if name.endswith("_s") else slice(None)
        got, want = mf.to_host(out)[rows], ref[name][rows]
        err = np.abs(got - want).max() / np.abs(want).max()
        print("  %s: |gpu-reference| %.2e" % (name, err), flush=True)
        assert err <= 1e-12, (name, err)


def test_derivative_rejects_bad_sizes():
    mf, torch = _mf()
    dc = mf.DerivativeCoefficients(o.grid_uniform(20), solver=False, so=6)
    x = mf.fortran_zeros((21, 4, 4))
    with pytest.raises(mf.MFError):
        mf.derivative(dc, x, 0, 1)


# ---------------------------------------------------------------- solver ---

def _make_problem(shape, bct, sigma, seed, grids=None):
    """Random interior u and bc; g = operator(u) so the discrete answer is ~u."""
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


def _solve_gpu(mf, grids, bct, g, bcs, sigma, precision, u0=None, tol=1e-12):
    sc = mf.SolverCoeffs3D(*grids, bct, 6)
    gd = mf.to_device(g)
    u0 = np.full(g.shape, 7.0, order="F") if u0 is None else u0
    ud = mf.to_device(u0)
    bc = mf.BoundaryCondition(*[mf.to_device(b) for b in bcs])
    rep = sc.solve(gd, ud, bc, sigma, precision, tol=tol)
    return mf.to_host(ud), rep


def _true_residual(sc, got, g, bcs, sigma):
    """||P(g - L u)||_2 / ||P g_f||_2 in 80-bit arithmetic."""
    ui = got[1:-1, 1:-1, 1:-1]
    gi = g[1:-1, 1:-1, 1:-1]
    r = o.multiplied_residual(sc, ui, gi, sigma, *bcs)
    b = o.multiplied_residual(sc, 0 * ui, gi, sigma, *bcs)
    return float(np.sqrt((r.astype(np.float64) ** 2).sum() / (b.astype(np.float64) ** 2).sum()))


CASES = [
    ((11, 12, 10), (1,) * 6, 2.0),
    ((11, 12, 10), (2,) * 6, 5.0),
    ((11, 12, 10), (1, 2, 2, 1, 1, 2), -40.0),
    ((34, 29, 40), (1,) * 6, 1e3),
    ((34, 29, 40), (1, 1, 2, 2, 1, 1), -300.0),
    ((66, 70, 21), (1,) * 6, 0.0),
    # 148 and 200 interior z points (k padded to 192 and 256): the fp16x2
    # solve runs its z pair as one kernel
    ((12, 14, 150), (1, 2, 1, 1, 2, 1), 50.0),
    ((10, 12, 202), (2, 1, 1, 2, 1, 1), -20.0),
]


def _exact_full(sc, g, bcs, sigma, shape):
    """(FP64 oracle solution, exact interior with the oracle's edges)."""
    ue = o.solve_exact(sc, g, *bcs, sigma)
    ref = np.full(shape, 7.0, order="F")
    o.solver_3d(sc, g, ref, *bcs, sigma)
    exact = ref.copy()
    exact[1:-1, 1:-1, 1:-1] = np.asarray(ue, dtype=np.float64)
    return ref, exact


def _faces_from(sc, u, bcs):
    """Re-apply the face rule (class_solver_3d.f90:261-297) to u's interior."""
    w = u.copy()
    cx, cy, cz = sc.cx, sc.cy, sc.cz
    bcx, bcy, bcz = bcs
    for c, bc, axis in ((cx, bcx, 0), (cy, bcy, 1), (cz, bcz, 2)):
        wl = np.moveaxis(w, axis, 0)
        inner = [slice(1, -1), slice(1, -1)]
        if c.bc[0] == 1:
            wl[0][tuple(inner)] = bc[:, :, 0]
        else:
            acc = bc[:, :, 0].copy()
            for m in range(1, 7):
                acc = acc - c.neum[0, m] * wl[m][tuple(inner)]
            wl[0][tuple(inner)] = acc * (1.0 / c.neum[0, 0])
        n = wl.shape[0] - 2
        if c.bc[1] == 1:
            wl[n + 1][tuple(inner)] = bc[:, :, 1]
        else:
            acc = bc[:, :, 1].copy()
            for m in range(5, -1, -1):
                acc = acc - c.neum[1, m] * wl[n - 5 + m][tuple(inner)]
            wl[n + 1][tuple(inner)] = acc * (1.0 / c.neum[1, 6])
    return w


@pytest.mark.parametrize("shape,bct,sigma", CASES)
def test_solve_fp64(shape, bct, sigma):
    mf, torch = _mf()
    grids, sc, g, bcs, uin = _make_problem(shape, bct, sigma, 11)
    tag = "fp64 shape=%s bct=%s sigma=%g" % (shape, bct, sigma)
    print("run " + tag, flush=True)
    got, _ = _solve_gpu(mf, grids, bct, g, bcs, sigma, mf.MF_FP64)
    ora, exact = _exact_full(sc, g, bcs, sigma, shape)
    exact = _faces_from(sc, exact, bcs)
    _check(tag, got, exact, ora)
    assert np.all(got[0, 0, :] == 7.0) and np.all(got[:, -1, -1] == 7.0)  # edges untouched


MIXED = {"fp32": 1, "bf16x3": 2, "fp16x2": 3}  # MF_MIXED, MF_MIXED_BF16X3, MF_MIXED_FP16X2


@pytest.mark.parametrize("inner", sorted(MIXED))
@pytest.mark.parametrize("shape,bct,sigma", CASES)
def test_solve_mixed(shape, bct, sigma, inner):
    """Default tol 1e-12: the residual contract; tol 1e-14: FP64-level error."""
    mf, torch = _mf()
    grids, sc, g, bcs, uin = _make_problem(shape, bct, sigma, 12)
    ora, exact = _exact_full(sc, g, bcs, sigma, shape)
    exact = _faces_from(sc, exact, bcs)
    for tol, err_bar in ((1e-12, 2e-11), (1e-14, 1e-12)):
        tag = "mixed-%s tol=%g shape=%s bct=%s sigma=%g" % (inner, tol, shape, bct, sigma)
        print("run " + tag, flush=True)
        got, rep = _solve_gpu(mf, grids, bct, g, bcs, sigma, MIXED[inner], tol=tol)
        true_res = _true_residual(sc, got, g, bcs, sigma)
        print("  refinement: %d inner solves, residual %.2e (first %.2e), 80-bit residual %.2e"
              % (rep.iterations, rep.residual, rep.residual0, true_res), flush=True)
        assert rep.residual <= tol and rep.iterations >= 2, (rep.iterations, rep.residual)
        assert true_res <= 2 * tol
        # at tol 1e-14 the error is FP64-level: within the bar of the FP64
        # solve (2x the FP64 oracle's own error)
        _check(tag, got, exact, ora, tol=err_bar, allow_oracle=(tol < 1e-13), ora_factor=2.0)
        assert np.all(got[0, 0, :] == 7.0) and np.all(got[:, -1, -1] == 7.0)


@pytest.mark.parametrize("shape", [(70, 29, 40), (180, 14, 12)])
@pytest.mark.parametrize("cz", ["0", "8", "13", "64"])
def test_solve_mixed_residual_kernels(monkeypatch, cz, shape):
    """Both FP64 residual kernels hold the refinement contract: the plane
    pipeline (MF_RESIDUAL_CZ=0) and the z-streaming kernel with 8 and 13
    output planes per block (several z chunks, a ragged last one) and with
    one chunk; 178 interior columns give x-tiles away from both x faces (the
    kernel's lane-shift x-stencils), 68 only tiles at a face.  Checked by
    the 80-bit residual of the returned solution."""
    monkeypatch.setenv("MF_RESIDUAL_CZ", cz)
    mf, torch = _mf()
    bct, sigma = (1, 1, 2, 2, 1, 1), -300.0
    grids, sc, g, bcs, uin = _make_problem(shape, bct, sigma, 13)
    for inner in ("fp16x2", "fp32"):
        print("run residual kernel cz=%s inner=%s shape=%s" % (cz, inner, shape), flush=True)
        got, rep = _solve_gpu(mf, grids, bct, g, bcs, sigma, MIXED[inner], tol=1e-13)
        true_res = _true_residual(sc, got, g, bcs, sigma)
        print("  %d inner solves, residual %.2e, 80-bit residual %.2e"
              % (rep.iterations, rep.residual, true_res), flush=True)
        assert rep.residual <= 1e-13 and true_res <= 2e-13


@pytest.mark.parametrize("shape,switch", [((140, 200, 136), "MF_H2_PLANE"),
                                          ((258, 150, 130), "MF_H2_PLANE"),
                                          ((140, 200, 136), "MF_H2_COLS"),
                                          ((136, 70, 200), "MF_H2_COLS")])
def test_solve_fp16x2_plane_pairs(monkeypatch, shape, switch):
    """The fp16x2 x/y transform pairs per z-plane (MF_H2_PLANE; x and y of
    129..256 points, k padded to 192 and to 256; 134 planes with the fused z
    pair, 128 planes with separate z transforms) and the column-block z pair
    (MF_H2_COLS; 134 and 198 z points, k padded to 192 and 256, the last
    column block partial) against the kernels they replace (switch = 0) and
    the FP64 solve: the residual contract checked by the 80-bit residual,
    2e-11 against FP64 each, and the two paths differing (so that both ran)."""
    mf, torch = _mf()
    bct, sigma = (1, 2, 1, 1, 2, 1), -300.0
    nx, ny, nz = shape
    grids = (o.grid_uniform(nx, 0, 1), o.grid_tanh(ny, 0, 1, 1.5), o.grid_asin(nz, 0, 0.5))
    sc = o.SolverCoeffs3D(*grids, bct, 6)
    rng = np.random.default_rng(14)
    g = np.asfortranarray(rng.standard_normal(shape))
    bcs = tuple(np.asfortranarray(rng.standard_normal(s))
                for s in ((ny - 2, nz - 2, 2), (nx - 2, nz - 2, 2), (nx - 2, ny - 2, 2)))
    print("run fp64 shape=%s" % (shape,), flush=True)
    u64, _ = _solve_gpu(mf, grids, bct, g, bcs, sigma, mf.MF_FP64)
    inner = (slice(1, -1),) * 3
    scale = np.abs(u64[inner]).max()
    got = {}
    for plane in ("1", "0"):
        monkeypatch.setenv(switch, plane)
        print("run fp16x2 %s=%s shape=%s" % (switch, plane, shape), flush=True)
        u, rep = _solve_gpu(mf, grids, bct, g, bcs, sigma, MIXED["fp16x2"])
        res = _true_residual(sc, u, g, bcs, sigma)
        err = float(np.abs(u[inner] - u64[inner]).max() / scale)
        print("  %d inner solves, residual %.2e, 80-bit residual %.2e, vs fp64 %.2e"
              % (rep.iterations, rep.residual, res, err), flush=True)
        assert rep.residual <= 1e-12 and res <= 2e-12 and err <= 2e-11
        got[plane] = u
    diff = float(np.abs(got["1"][inner] - got["0"][inner]).max() / scale)
    print("  %s=1 vs 0: %.2e" % (switch, diff), flush=True)
    # both within 2e-11 of FP64, so within 4e-11 of each other; nonzero:
    # both kernels ran
    assert 0.0 < diff <= 4e-11


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
        ora = np.zeros_like(u)
        o.solver_3d(sc, g, ora, bcx, bcy, bcz, sigma)
        ue = o.solve_exact(sc, g, bcx, bcy, bcz, sigma)
        exact = ora.copy()
        exact[1:-1, 1:-1, 1:-1] = np.asarray(ue, dtype=np.float64)
        for prec in (mf.MF_FP64, mf.MF_MIXED, mf.MF_MIXED_BF16X3, mf.MF_MIXED_FP16X2):
            tag = "c1 sigma=%g precision=%s" % (sigma, ("fp64", "mixed", "bf16x3", "fp16x2")[prec])
            print("run " + tag, flush=True)
            got, rep = _solve_gpu(mf, (gx, gy, gz), bct, g, (bcx, bcy, bcz), sigma, prec,
                                  u0=np.zeros_like(u))
            _check(tag, got, exact, ora, allow_oracle=(prec == 0))
            inner = (slice(1, -1),) * 3
            # the manufactured u is recovered to the conditioning of g
            assert np.abs(got[inner] - u[inner]).max() <= 1e-10


def test_all_neumann_poisson_incompatible_rhs():
    """sigma = 0, Neumann on every face, a right-hand side with a large
    incompatible part (an offset 100 times its compatible part) on a
    tanh-2.5 grid whose null eigenvalue LAPACK returns near 1e-11: the
    device (its own eigen-decomposition) and the oracle (LAPACK's) agree
    to 1e-11, constant included, because both remove the null component
    before the transforms."""
    mf, torch = _mf()
    shape = (40, 64, 36)
    gs = (o.grid_uniform(shape[0], 0, 2 * np.pi), o.grid_tanh(shape[1], -1, 1, 2.5),
          o.grid_uniform(shape[2], 0, np.pi))
    X, Y, Z = np.meshgrid(*gs, indexing="ij")
    g = np.asfortranarray(np.sin(X) * np.cos(np.pi * Y) * np.cos(2 * Z) + 100.0)
    bcs = [np.zeros(s, order="F") for s in ((62, 34, 2), (38, 34, 2), (38, 62, 2))]
    sc_o = o.SolverCoeffs3D(*gs, (2,) * 6, 6)
    ref = np.zeros(shape, order="F")
    o.solver_3d(sc_o, g, ref, *bcs, 0.0)
    sc = mf.SolverCoeffs3D(*gs, (2,) * 6, 6)
    u = mf.fortran_zeros(shape)
    print("run all-neumann poisson %s" % (shape,), flush=True)
    sc.solve(mf.to_device(g), u, mf.BoundaryCondition(*[mf.to_device(b) for b in bcs]), 0.0, mf.MF_FP64)
    got = mf.to_host(u)
    inner = (slice(1, -1),) * 3
    err = np.abs(got - ref)[inner].max() / np.abs(ref[inner]).max()
    print("all-neumann poisson: |gpu-oracle64| %.2e" % err, flush=True)
    assert err <= 1e-11

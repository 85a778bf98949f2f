"""Pins the CPU oracle before anything is compared against it.

The reference holds no stored vectors for this path (its tests print errors of
manufactured solutions), so each oracle function is pinned by what the
mathematics fixes, chosen so that a plausible slip (dropped term, wrong sign,
wrong index, transposed operand) fails at least one check:

* der_coeffs_generic / der_coeffs_global: Lele's closed-form 6th-order
  interior coefficients on a uniform grid; polynomial exactness of every row
  (including the one-sided edge rows) on a stretched grid; order of accuracy.
* pentalu / pentares: against a dense LAPACK solve of the unfactored matrix.
* derivative_3d (all axes, both modes): exact derivatives of polynomials.
* solver coefficients: polynomial exactness of the 1-D solver operator with its
  Dirichlet / Neumann boundary columns; the eigen-decomposition identities.
* solve_3d / solver_3d: against a dense Kronecker-product solve on small grids
  (Dirichlet, Neumann, mixed), the reference's manufactured x^9+y^9+xy case
  (test_solver_3d.f90:32-257), and the Neumann extrapolation identity.
* RNG: SplitMix64 known-answer vector (seed 1234567).
"""
import numpy as np
import pytest

import mf_oracle as o


def test_splitmix64_known_answer():
    # published SplitMix64 outputs for state 1234567
    ref = [6457827717110365317, 3203168211198807973, 9817491932198370423,
           4593380528125082431, 16408922859458223821]
    assert [int(v) for v in o.splitmix64(1234567, 5)] == ref


@pytest.mark.parametrize("so,der,cl", [(8, 2, 3), (8, 2, 1), (6, 2, 2), (6, 1, 3), (4, 2, 1)])
def test_taylor_solve_extended_vs_mpmath(so, der, cl):
    """The extended-precision Taylor solve equals a 40-digit mpmath solve of
    the same FP64 offsets to ~1e-15, and the FP64 dgesv variant is within
    the system's condition number of it."""
    import mpmath

    g = o.grid_tanh(41, 0.0, 1.0, 1.6)
    i = {1: 0, 2: 1, 3: 20}[cl]
    h0 = np.zeros(7)
    if cl == 3:
        h0[0:5] = g[i] - g[i - 2:i + 3]
    else:
        h0[:] = g[i] - g[0:7]
    ci, ce = o.der_coeffs_generic(h0, cl, der, so)
    n1, n2 = o._DER_STENCIL[cl][so]
    nt = n1 + n2
    o.TAYLOR = "lapack"
    try:
        ci64, ce64 = o.der_coeffs_generic(h0, cl, der, so)
    finally:
        o.TAYLOR = "extended"
    # the same system in 40-digit arithmetic (offsets selected as the
    # reference does, class_derivatives_coefficient.f90:59-127)
    mpmath.mp.dps = 40
    h = np.zeros(10)
    if cl == 1:
        h[0:n1] = h0[1:n1 + 1]; h[n1:nt] = h0[0:n2]
    elif cl == 2:
        h[0] = h0[0]; h[1:n1] = h0[2:n1 + 1]; h[n1:nt] = h0[0:n2]
    else:
        s = (4 - n1) // 2
        h[0:2] = h0[s:2 + s]
        if n1 >= 3 - s:
            h[2 - s:n1] = h0[3:n1 + 1 + s]
        s = (5 - n2) // 2
        h[n1:nt] = h0[s:n2 + s]
    A = mpmath.zeros(nt, nt)
    for j in range(n1):
        A[der, j] = -1
        for r in range(der + 1, nt):
            A[r, j] = A[r - 1, j] * mpmath.mpf(h[j]) / (r - der)
    for j in range(n1, nt):
        A[0, j] = 1
        for r in range(1, nt):
            A[r, j] = A[r - 1, j] * mpmath.mpf(h[j]) / r
    rhs = mpmath.zeros(nt, 1)
    rhs[der] = 1
    x = mpmath.lu_solve(A, rhs)
    exact = np.array([float(x[m]) for m in range(nt)])
    s = (5 - n2) // 2 if cl == 3 else 0
    scale = np.abs(exact).max()
    np.testing.assert_allclose(ce[s:s + n2], exact[n1:nt], atol=4e-15 * scale, rtol=0)
    # FP64 dgesv differs by at most the conditioning of the system
    cond = float(mpmath.norm(A, 1) * mpmath.norm(mpmath.inverse(A), 1))
    assert np.abs(ce64 - ce).max() <= 10 * cond * 1.1e-16 * scale


def test_normal_moments():
    z = o.normal(0, 200001)
    assert abs(z.mean()) < 1e-2 and abs(z.std() - 1.0) < 1e-2


@pytest.mark.parametrize("n", [12, 33])
def test_lele_interior_closed_form(n):
    g = o.grid_uniform(n, 0.0, 2.0)
    h = g[1] - g[0]
    dl, dr = o.der_coeffs_global(g, 1, 6)
    a, b = 14.0 / 9.0, 1.0 / 9.0
    for i in range(2, n - 2):
        np.testing.assert_allclose(dl[i], [0, 1 / 3, 1, 1 / 3, 0], atol=1e-13)
        np.testing.assert_allclose(dr[i, :5] * h, [-b / 4, -a / 2, 0, a / 2, b / 4], atol=1e-12)
    ddl, ddr = o.der_coeffs_global(g, 2, 6)
    a, b = 12.0 / 11.0, 3.0 / 11.0
    for i in range(2, n - 2):
        np.testing.assert_allclose(ddl[i], [0, 2 / 11, 1, 2 / 11, 0], atol=1e-13)
        np.testing.assert_allclose(ddr[i, :5] * h * h,
                                   [b / 4, a, -2 * a - b / 2, a, b / 4], atol=1e-11)


def _dense_from_dl(dl):
    """Dense quasi-pentadiagonal matrix from the unfactored packed rows."""
    n = dl.shape[0]
    A = np.zeros((n, n))
    A[0, 0:4] = [dl[0, 2], dl[0, 3], dl[0, 4], dl[0, 0]]
    A[1, 0:4] = [dl[1, 1], dl[1, 2], dl[1, 3], dl[1, 4]]
    for i in range(2, n - 2):
        A[i, i - 2:i + 3] = dl[i]
    A[n - 2, n - 4:n] = dl[n - 2, 0:4]
    A[n - 1, n - 4] = dl[n - 1, 4]
    A[n - 1, n - 3:n] = dl[n - 1, 0:3]
    return A


def _dense_rhs(dr):
    n = dr.shape[0]
    R = np.zeros((n, n))
    R[0, 0:7] = dr[0]
    R[1, 0:7] = dr[1]
    for i in range(2, n - 2):
        R[i, i - 2:i + 3] = dr[i, 0:5]
    R[n - 2, n - 7:n] = dr[n - 2]
    R[n - 1, n - 7:n] = dr[n - 1]
    return R


@pytest.mark.parametrize("so", [4, 6, 8])
@pytest.mark.parametrize("der", [1, 2])
def test_pentalu_pentares_vs_dense(so, der):
    g = o.grid_tanh(40, 0.0, 1.0, 1.7)
    dl, _ = o.der_coeffs_global(g, der, so)
    lu = o.pentalu(dl)
    A = _dense_from_dl(dl)
    rng = np.random.default_rng(3)
    b = rng.standard_normal((40, 5))
    np.testing.assert_allclose(o.pentares(lu, b), np.linalg.solve(A, b), rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("so", [4, 6, 8])
@pytest.mark.parametrize("der", [1, 2])
def test_rows_polynomial_exact(so, der):
    """Every row of the compact scheme is exact for polynomials of degree <= nt-1."""
    g = o.grid_tanh(30, -0.5, 0.5, 1.5)
    dl, dr = o.der_coeffs_global(g, der, so)
    A = _dense_from_dl(dl)
    R = _dense_rhs(dr)
    deg = {4: 4, 6: 6, 8: 8}[so]
    # minimal row degree: cl=1 (n1+n2-1), cl=2, cl=3 from the stencil table
    deg = min(sum(o._DER_STENCIL[cl][so]) - 1 for cl in (1, 2, 3))
    for p in range(deg + 1):
        f = g ** p
        fd = p * g ** (p - 1) if der == 1 else p * (p - 1) * g ** max(p - 2, 0)
        if p == 0:
            fd = np.zeros_like(g)
        if der == 2 and p == 1:
            fd = np.zeros_like(g)
        np.testing.assert_allclose(A @ fd, R @ f, atol=1e-9 * max(1.0, np.abs(R @ f).max()))
    # one degree above the interior order must fail (the check has teeth)
    p = deg + 1
    f = g ** p
    fd = p * g ** (p - 1) if der == 1 else p * (p - 1) * g ** (p - 2)
    assert np.abs(A @ fd - R @ f).max() > 1e-9


@pytest.mark.parametrize("order", [1, 2])
def test_order_of_accuracy(order):
    errs = []
    ns = [33, 65, 129]
    for n in ns:
        g = o.grid_tanh(n, 0.0, 1.0, 1.2)
        dc = o.DerCoeffs(g, solver=False, so=6)
        f = np.sin(2 * np.pi * g) + np.cos(4 * np.pi * g)
        if order == 1:
            ex = 2 * np.pi * np.cos(2 * np.pi * g) - 4 * np.pi * np.sin(4 * np.pi * g)
            d = dc.der(f)
        else:
            ex = -(2 * np.pi) ** 2 * np.sin(2 * np.pi * g) - (4 * np.pi) ** 2 * np.cos(4 * np.pi * g)
            d = dc.dder(f)
        errs.append(np.abs(d - ex).max() / np.abs(ex).max())
    rates = [np.log(errs[i] / errs[i + 1]) / np.log(2.0) for i in range(2)]
    # boundary closures are 5th/4th order (one-sided), interior 6th
    assert min(rates) > 3.5, (errs, rates)


def _poly3d(gx, gy, gz, coef):
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    f = np.zeros_like(X)
    dfs = [np.zeros_like(X) for _ in range(6)]
    for (a, b, c), w in coef.items():
        f += w * X ** a * Y ** b * Z ** c
        dfs[0] += w * a * X ** max(a - 1, 0) * Y ** b * Z ** c
        dfs[1] += w * a * (a - 1) * X ** max(a - 2, 0) * Y ** b * Z ** c
        dfs[2] += w * b * X ** a * Y ** max(b - 1, 0) * Z ** c
        dfs[3] += w * b * (b - 1) * X ** a * Y ** max(b - 2, 0) * Z ** c
        dfs[4] += w * c * X ** a * Y ** b * Z ** max(c - 1, 0)
        dfs[5] += w * c * (c - 1) * X ** a * Y ** b * Z ** max(c - 2, 0)
    return np.asfortranarray(f), [np.asfortranarray(d) for d in dfs]


@pytest.mark.parametrize("solver", [False, True])
def test_derivative_3d_polynomial(solver):
    gx = o.grid_uniform(14, 0.0, 1.0)
    gy = o.grid_tanh(17, 0.0, 1.0, 1.5)
    gz = o.grid_asin(12, 0.0, 0.5)
    coef = {(5, 2, 1): 1.0, (1, 6, 0): -0.5, (2, 3, 6): 0.25, (0, 0, 0): 2.0}
    f, dfs = _poly3d(gx, gy, gz, coef)
    dcs = [o.DerCoeffs(g, solver=solver, so=6) for g in (gx, gy, gz)]
    for axis in range(3):
        for order in (1, 2):
            d = o.derivative_3d(dcs[axis], f, axis, order)
            ex = dfs[2 * axis + order - 1]
            if solver:
                # points on >= 2 boundary planes (box edges) are undefined in
                # the reference: boundary-plane lines skip their two ends and
                # the lines lying in two boundary planes are never written
                nb = sum(((np.arange(s) == 0) | (np.arange(s) == s - 1)).reshape(
                    [-1 if q == a else 1 for q in range(3)]).astype(int)
                    for a, s in enumerate(f.shape))
                m = nb <= 1
                np.testing.assert_allclose(d[m], ex[m], atol=2e-7 * np.abs(ex).max())
                assert np.all(d[~m] == 0.0)
            else:
                np.testing.assert_allclose(d, ex, atol=2e-7 * np.abs(ex).max())


# ---------------------------------------------------------------- solver ----

def _grid(n, kind):
    if kind == "u":
        return o.grid_uniform(n, 0.0, 1.0)
    if kind == "t":
        return o.grid_tanh(n, 0.0, 1.0, 1.5)
    return o.grid_asin(n, -0.5, 0.5)


@pytest.mark.parametrize("bct", [(1, 1), (2, 2), (1, 2), (2, 1)])
@pytest.mark.parametrize("kind", ["u", "t", "a"])
def test_solver_1d_polynomial_exact(bct, kind):
    """A u + inv(lhs)(bcol.bc) equals u'' for polynomials up to the boundary rows' degree."""
    g = _grid(24, kind)
    c = o.SolverCoeffs1D(g, bct, 6)
    n = c.n
    for p in range(0, 6):
        u = g ** p
        upp = p * (p - 1) * g ** max(p - 2, 0)
        up = p * g ** max(p - 1, 0)
        bc1 = u[0] if bct[0] == 1 else up[0]
        bc2 = u[-1] if bct[1] == 1 else up[-1]
        t = c.rhs @ u[1:-1]
        t[0] += c.ddr[0, 0] * bc1
        t[1] += c.ddr[1, 0] * bc1
        t[n - 2] += c.ddr[n - 2, 7] * bc2
        t[n - 1] += c.ddr[n - 1, 6] * bc2
        np.testing.assert_allclose(c.inv_lhs @ t, upp[1:-1], atol=1e-8 * max(1, np.abs(upp).max()))
    # Neumann extrapolation weights: exact first derivative of degree <= 6
    for p in range(0, 7):
        u = g ** p
        up = p * g ** max(p - 1, 0)
        np.testing.assert_allclose(c.neum[0] @ u[0:7], up[0], atol=1e-7 * max(1, abs(up[0])))
        np.testing.assert_allclose(c.neum[1] @ u[-7:], up[-1], atol=1e-7 * max(1, abs(up[-1])))


@pytest.mark.parametrize("kind", ["u", "t", "a"])
def test_solver_eigen_identities(kind):
    c = o.SolverCoeffs1D(_grid(30, kind), (1, 1), 6)
    A = c.operator()
    np.testing.assert_allclose(A @ c.vect, c.vect * c.eigen[None, :], atol=1e-8 * np.abs(c.eigen).max())
    np.testing.assert_allclose(c.inv_vect @ c.vect, np.eye(c.n), atol=1e-10)
    np.testing.assert_allclose(c.inv_lhs @ c.lhs, np.eye(c.n), atol=1e-12)
    assert np.all(c.eigen < 0)  # Dirichlet second derivative: negative spectrum


def _kron_operator(sc, sigma):
    """Dense operator on the Fortran-ordered interior vector (i fastest)."""
    Ax, Ay, Az = sc.cx.operator(), sc.cy.operator(), sc.cz.operator()
    Ix, Iy, Iz = np.eye(Ax.shape[0]), np.eye(Ay.shape[0]), np.eye(Az.shape[0])
    K = (np.kron(Iz, np.kron(Iy, Ax)) + np.kron(Iz, np.kron(Ay, Ix))
         + np.kron(Az, np.kron(Iy, Ix)))
    return K + sigma * np.eye(K.shape[0])


def _random_problem(shape, bct, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    bcx = rng.standard_normal((ny - 2, nz - 2, 2))
    bcy = rng.standard_normal((nx - 2, nz - 2, 2))
    bcz = rng.standard_normal((nx - 2, ny - 2, 2))
    g = np.asfortranarray(rng.standard_normal(shape))
    return g, np.asfortranarray(bcx), np.asfortranarray(bcy), np.asfortranarray(bcz)


@pytest.mark.parametrize("bct,sigma", [((1,) * 6, 2.0), ((1,) * 6, -300.0),
                                       ((2,) * 6, 5.0), ((1, 2, 2, 1, 1, 2), 0.0),
                                       ((1, 1, 2, 2, 1, 1), 1e3)])
def test_solve_vs_dense_kronecker(bct, sigma):
    shape = (11, 12, 10)
    gx = o.grid_uniform(shape[0], 0.0, 1.0)
    gy = o.grid_tanh(shape[1], 0.0, 1.0, 1.5)
    gz = o.grid_asin(shape[2], 0.0, 0.5)
    sc = o.SolverCoeffs3D(gx, gy, gz, bct, 6)
    g, bcx, bcy, bcz = _random_problem(shape, bct, 7)
    u = np.asfortranarray(np.full(shape, 123.0))
    o.solver_3d(sc, g, u, bcx, bcy, bcz, sigma)
    # dense route: K u = g_int - (BC part of the operator applied to zero field)
    zero = np.zeros((shape[0] - 2, shape[1] - 2, shape[2] - 2))
    bpart = o.apply_operator(sc, zero, 0.0, bcx, bcy, bcz)
    K = _kron_operator(sc, sigma)
    rhs = (g[1:-1, 1:-1, 1:-1] - bpart).ravel(order="F")
    ud = np.linalg.solve(K, rhs).reshape(zero.shape, order="F")
    np.testing.assert_allclose(u[1:-1, 1:-1, 1:-1], ud, rtol=0, atol=1e-10 * np.abs(ud).max())
    # residual with the oracle's operator
    res = o.apply_operator(sc, u[1:-1, 1:-1, 1:-1], sigma, bcx, bcy, bcz) - g[1:-1, 1:-1, 1:-1]
    assert np.abs(res).max() < 1e-9 * np.abs(g).max()
    # faces: Dirichlet copies bc, Neumann satisfies the one-sided derivative = bc
    if bct[0] == 1:
        assert np.array_equal(u[0, 1:-1, 1:-1], bcx[:, :, 0])
    else:
        d = np.tensordot(sc.cx.neum[0], u[0:7, 1:-1, 1:-1], axes=(0, 0))
        np.testing.assert_allclose(d, bcx[:, :, 0], atol=1e-9)
    if bct[3] == 2:
        d = np.tensordot(sc.cy.neum[1], u[1:-1, -7:, 1:-1], axes=(0, 1))
        np.testing.assert_allclose(d, bcy[:, :, 1], atol=1e-9)
    # edges untouched
    assert u[0, 0, 0] == 123.0 and u[-1, -1, 5] == 123.0 and u[4, 0, 0] == 123.0


def test_manufactured_x9_y9_xy():
    """test_solver_3d.f90:32-257: u = x^9+y^9+xy, sigma=2, all Dirichlet."""
    n = 33
    gx, gy, gz = (o.grid_asin(n, -0.5, 0.5), o.grid_asin(n, -0.5, 0.5), o.grid_asin(n, -0.25, 0.25))
    sc = o.SolverCoeffs3D(gx, gy, gz, (1,) * 6, 6)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    uex = X ** 9 + Y ** 9 + X * Y
    g = np.asfortranarray(72 * X ** 7 + 72 * Y ** 7 + 2.0 * uex)
    bcx = np.asfortranarray(np.stack([uex[0, 1:-1, 1:-1], uex[-1, 1:-1, 1:-1]], axis=-1))
    bcy = np.asfortranarray(np.stack([uex[1:-1, 0, 1:-1], uex[1:-1, -1, 1:-1]], axis=-1))
    bcz = np.asfortranarray(np.stack([uex[1:-1, 1:-1, 0], uex[1:-1, 1:-1, -1]], axis=-1))
    u = np.asfortranarray(np.zeros_like(uex))
    o.solver_3d(sc, g, u, bcx, bcy, bcz, 2.0)
    err = np.sqrt(((u - uex)[1:-1, 1:-1, 1:-1] ** 2).sum() / (uex[1:-1, 1:-1, 1:-1] ** 2).sum())
    assert err < 1e-6, err


@pytest.mark.parametrize("face", [0, 3, 5])
def test_influence_impulses_vs_dense(face):
    """md_influence_matrix_set restated, against a dense Kronecker solve and
    dense derivative rows of the same discretisation."""
    shape = (10, 11, 9)
    bct = (1, 1, 1, 1, 1, 2) if face != 5 else (1, 2, 1, 1, 1, 1)
    gs = (o.grid_uniform(10), o.grid_tanh(11, 0, 1, 1.3), o.grid_asin(9, 0, 1))
    sc = o.SolverCoeffs3D(*gs, bct, 6)
    dcs = [o.DerCoeffs(g, solver=True, so=6) for g in gs]
    sigma = -50.0
    got = o.influence_impulses(sc, dcs, sigma, face, 2, 4)
    n1, n2, n3 = 8, 9, 7
    K = _kron_operator(sc, sigma)
    D = []
    for g, dc in zip(gs, dcs):
        n = len(g)
        E = np.eye(n)
        D.append(np.stack([dc.der(E[:, m]) for m in range(n)], 1))  # D[:, m] = der(e_m)
    for b, t in enumerate(range(2, 6)):
        bcs = [np.zeros((n2, n3, 2)), np.zeros((n1, n3, 2)), np.zeros((n1, n2, 2))]
        l, m = face // 2, face % 2
        L1 = n2 if l == 0 else n1
        bcs[l][t % L1, t // L1, m] = 1.0
        zero = np.zeros((n1, n2, n3))
        bpart = o.apply_operator(sc, zero, 0.0, *bcs)
        ui = np.linalg.solve(K, -bpart.ravel(order="F")).reshape((n1, n2, n3), order="F")
        u = np.zeros(shape)
        u[1:-1, 1:-1, 1:-1] = ui
        for ax, c in enumerate((sc.cx, sc.cy, sc.cz)):   # faces as the solver writes them
            ul = np.moveaxis(u, ax, 0)
            inner = (slice(1, -1), slice(1, -1))
            for side in (0, 1):
                idx = 0 if side == 0 else ul.shape[0] - 1
                if c.bc[side] == 1:
                    ul[idx][inner] = bcs[ax][:, :, side]
                elif side == 0:
                    ul[0][inner] = -sum(c.neum[0, k] * ul[k][inner] for k in range(1, 7)) / c.neum[0, 0]
                else:
                    nn = ul.shape[0] - 2
                    ul[nn + 1][inner] = -sum(c.neum[1, k] * ul[nn - 5 + k][inner]
                                             for k in range(6)) / c.neum[1, 6]
        dx = np.tensordot(D[0], u, axes=(1, 0))
        dy = np.moveaxis(np.tensordot(D[1], u, axes=(1, 1)), 0, 1)
        dz = np.moveaxis(np.tensordot(D[2], u, axes=(1, 2)), 0, 2)
        ref = np.concatenate([np.stack([dx[0, 1:-1, 1:-1], -dx[-1, 1:-1, 1:-1]], -1).ravel(order="F"),
                              np.stack([dy[1:-1, 0, 1:-1], -dy[1:-1, -1, 1:-1]], -1).ravel(order="F"),
                              np.stack([dz[1:-1, 1:-1, 0], -dz[1:-1, 1:-1, -1]], -1).ravel(order="F")])
        np.testing.assert_allclose(got[b], ref, atol=1e-9 * np.abs(ref).max())


def test_poisson_all_neumann_null_mode():
    shape = (10, 11, 9)
    sc = o.SolverCoeffs3D(o.grid_uniform(10), o.grid_tanh(11, 0, 1, 1.2), o.grid_uniform(9), (2,) * 6, 6)
    # one eigenvalue per axis is (numerically) zero
    for c in (sc.cx, sc.cy, sc.cz):
        assert np.sum(np.abs(c.eigen) < 1e-9) == 1
    rng = np.random.default_rng(1)
    uin = rng.standard_normal((8, 9, 7))
    g = np.zeros(shape, order="F")
    g[1:-1, 1:-1, 1:-1] = o.apply_operator(sc, uin, 0.0)
    z = lambda a, b: np.zeros((a, b, 2), order="F")  # noqa: E731
    u = np.zeros(shape, order="F")
    o.solver_3d(sc, g, u, z(9, 7), z(8, 7), z(8, 9), 0.0)
    # solution equals uin up to the null vector (a constant for this operator)
    diff = u[1:-1, 1:-1, 1:-1] - uin
    assert np.abs(diff - diff.mean()).max() < 1e-8


def test_null_eigenvalue_snapped_to_zero():
    """The constant null vector of an axis with Neumann conditions at both
    ends gets eigenvalue exactly 0 (the reference's absolute 1e-11 guard,
    class_solver_3d.f90:207-209, misses dgeev's 1.05e-11 on this 256-point
    tanh-2.5 axis); no other eigenvalue changes, Dirichlet axes are untouched."""
    g = o.grid_tanh(256, -1.0, 1.0, 2.5)
    for bct, nzero in (((2, 2), 1), ((1, 1), 0), ((1, 2), 0)):
        c = o.SolverCoeffs1D(g, bct, 6)
        lhs, rhs = o.solver_lhs_rhs(c.ddl, c.ddr)
        raw = o.solver_matrix_init(c.ddl, c.ddr)[0]
        assert np.sum(c.eigen == 0.0) == nzero
        changed = c.eigen != raw
        assert np.sum(changed) == nzero
        if nzero:
            # the snapped one is the null mode: its eigenvector is constant
            i = int(np.argmin(np.abs(raw)))
            assert c.eigen[i] == 0.0 and abs(raw[i]) > 1e-11 * 0.5
            v = c.vect[:, i]
            assert np.abs(v - v.mean()).max() < 1e-8 * np.abs(v).max()


def test_null_component_removal_all_neumann():
    """sigma = 0 with Neumann faces everywhere: the solve drops c0 = w0.g /
    w0.1 (w0 the left null vector) from the folded right-hand side.  Since
    w0^T L = 0, for g = L u_in + 5 the removed part is exactly the offset 5,
    the solution is u_in up to a constant, and L u equals the compatible part
    g - c0."""
    shape = (14, 16, 12)
    gs = (o.grid_uniform(shape[0], 0, 2), o.grid_tanh(shape[1], -1, 1, 2.0), o.grid_asin(shape[2], 0, 1))
    sc = o.SolverCoeffs3D(*gs, (2,) * 6, 6)
    rows = o.null_rows(sc)
    assert rows is not None
    rng = np.random.default_rng(4)
    uin = rng.standard_normal(tuple(s - 2 for s in shape))
    gi = o.apply_operator(sc, uin, 0.0) + 5.0
    wx, wy, wz = rows
    c0 = np.einsum("i,j,k,ijk->", wx, wy, wz, gi) / (wx.sum() * wy.sum() * wz.sum())
    assert abs(c0 - 5.0) < 1e-9
    g = np.zeros(shape, order="F")
    g[1:-1, 1:-1, 1:-1] = gi
    z = [np.zeros(s, order="F") for s in ((14, 10, 2), (12, 10, 2), (12, 14, 2))]
    u = np.zeros(shape, order="F")
    o.solver_3d(sc, g, u, *z, 0.0)
    ui = u[1:-1, 1:-1, 1:-1]
    d = ui - uin
    assert np.abs(d - d.mean()).max() < 1e-9 * np.abs(uin).max()
    res = o.apply_operator(sc, ui, 0.0) - (gi - c0)
    assert np.abs(res).max() < 1e-9 * np.abs(gi).max()
    # a problem that is not singular (sigma = 0 but a Dirichlet axis) is untouched
    sc2 = o.SolverCoeffs3D(*gs, (2, 2, 1, 1, 2, 2), 6)
    assert o.null_rows(sc2) is None
    r = rng.standard_normal(uin.shape)
    assert o.remove_null_component(sc2, r) is r

"""CPU oracle: FP64 restatement of MFLOPS3D-MD's compact derivatives and
matrix-diagonalisation Helmholtz/Poisson solver.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path never calls it.

Every routine follows the reference Fortran step by step; the file:line it
restates is cited in its docstring (paths relative to the reference root).
LAPACK/BLAS calls of the reference map to numpy's LAPACK/BLAS:
``dgesv`` -> ``numpy.linalg.solve``, ``dgetrf+dgetri`` -> ``numpy.linalg.inv``,
``dgeev`` -> ``numpy.linalg.eig``, ``dgemm`` -> ``@``.

Array convention: 3-D fields are numpy arrays of shape (nx, ny, nz) in
Fortran order, so ``f[i, j, k]`` is the reference's ``f(i+1, j+1, k+1)`` and
the memory layout is the reference's column-major one.  Line coefficient
tables keep the reference's column order: ``dl[:, 0:5] = (p, f, c, d, e)``.

Parity pinning (what independently checks each part) is listed in
``tests/test_oracle.py``; no function here is left unpinned.
"""
from __future__ import annotations

import numpy as np

# ----------------------------------------------------------------------------
# Grids  (src/class_mesh/class_mesh.f90:112-335)
# ----------------------------------------------------------------------------


def grid_uniform(n: int, a: float = 0.0, b: float = 1.0) -> np.ndarray:
    """Uniform grid of n points on [a, b] (class_mesh.f90:164,178 xi = xa+(i-1)*pas)."""
    pas = (b - a) / (n - 1)
    return a + np.arange(n, dtype=np.float64) * pas


def grid_tanh(n: int, a: float, b: float, alpha: float) -> np.ndarray:
    """tanh stretching x = xa+L(1+tanh(alpha(2 xi-1))/tanh alpha)/2.

    class_mesh.f90:203 (present in the reference only as a commented line);
    xi is the uniform coordinate scaled to [0,1].
    """
    s = np.arange(n, dtype=np.float64) / (n - 1)
    return a + (b - a) * (1.0 + np.tanh(alpha * (2.0 * s - 1.0)) / np.tanh(alpha)) / 2.0


def grid_asin(n: int, a: float, b: float, beta: float = 33.3) -> np.ndarray:
    """The reference's live asin/cos stretching (class_mesh.f90:181-200)."""
    alpha = 1.0 - beta / float(n * n)
    pas = (b - a) / (n - 1)
    xi = a + np.arange(n, dtype=np.float64) * pas
    return a + (b - a) * (np.arcsin(-alpha * np.cos(np.pi * (xi - a) / (b - a)))
                          / np.arcsin(alpha) + 1.0) * 0.5


# ----------------------------------------------------------------------------
# Seeded synthetic inputs: SplitMix64 -> uniform [0,1) -> Box-Muller normal.
# The same generator is implemented in the product's host code; both are
# defined by the published SplitMix64 constants (Steele, Lea, Flood 2014).
# ----------------------------------------------------------------------------

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """count consecutive SplitMix64 outputs for state=seed (uint64 array)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    """Uniform doubles in [0,1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, count: int) -> np.ndarray:
    """Standard normals by Box-Muller on consecutive uniform pairs."""
    m = (count + 1) // 2
    u = uniform01(seed, 2 * m)
    u1 = 1.0 - u[0::2]  # (0,1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:count]


# ----------------------------------------------------------------------------
# Derivative coefficients (src/class_derivatives/class_derivatives_coefficient.f90)
# ----------------------------------------------------------------------------

# (n1 implicit, n2 explicit) per (cl, so): class_derivatives_coefficient.f90:59-127
_DER_STENCIL = {
    1: {8: (3, 6), 6: (1, 6), 4: (1, 4), 2: (0, 3)},
    2: {8: (1, 5), 6: (2, 6), 4: (1, 5), 2: (0, 3)},
    3: {8: (4, 5), 6: (2, 5), 4: (2, 3), 2: (0, 3)},
}


# How the small Taylor systems are solved.  "extended": built and solved in
# x87 80-bit extended precision (numpy.longdouble) by Gaussian elimination
# with partial pivoting, then rounded to FP64 -- the FP64 LU of the reference
# (dgesv) spreads by up to ~5e-11 relative across LAPACK builds for the
# 8th-order second derivative, more than the 1e-12 parity bar, while the
# extended solve is accurate to ~1e-15.  "lapack": the reference's FP64 dgesv.
TAYLOR = "extended"
_XR = np.longdouble


def _taylor_matrix(h: np.ndarray, n1: int, nt: int, der: int) -> np.ndarray:
    """Taylor matrix a(nt,nt) (class_derivatives_coefficient.f90:135-151)."""
    dt = _XR if TAYLOR == "extended" else np.float64
    a = np.zeros((nt, nt), dtype=dt)
    hh = np.asarray(h, dtype=np.float64).astype(dt)
    # implicit columns: a(der+1,j)=-1, a(i,j)=a(i-1,j)*h(j)/(i-der-1)
    a[der, :n1] = -1
    for i in range(der + 1, nt):  # Fortran i = der+2..nt
        for j in range(n1):
            a[i, j] = a[i - 1, j] * hh[j] / dt(i - der)
    # explicit columns: a(1,j)=1, a(i,j)=a(i-1,j)*h(j)/(i-1)
    a[0, n1:nt] = 1
    for i in range(1, nt):
        for j in range(n1, nt):
            a[i, j] = a[i - 1, j] * hh[j] / dt(i)
    return a


def _taylor_solve(a: np.ndarray, r: np.ndarray) -> np.ndarray:
    """Solve the Taylor system (dgesv at class_derivatives_coefficient.f90:159,
    class_solver_coefficient_3d.f90:212); see TAYLOR."""
    if TAYLOR != "extended":
        return np.linalg.solve(np.asarray(a, np.float64), np.asarray(r, np.float64))
    a = np.array(a, dtype=_XR)
    r = np.array(r, dtype=_XR)
    nt = a.shape[0]
    for k in range(nt):
        col = np.abs(a[k:, k])
        p = k + int(np.argmax(col))  # first maximal pivot
        if col.max() == 0:
            raise np.linalg.LinAlgError("singular Taylor system")
        if p != k:
            a[[k, p]] = a[[p, k]]
            r[[k, p]] = r[[p, k]]
        for i in range(k + 1, nt):
            l = a[i, k] / a[k, k]
            a[i, k:] = a[i, k:] - l * a[k, k:]
            r[i] = r[i] - l * r[k]
    for i in range(nt - 1, -1, -1):
        s = r[i]
        for j in range(i + 1, nt):
            s = s - a[i, j] * r[j]
        r[i] = s / a[i, i]
    return r.astype(np.float64)


def der_coeffs_generic(h0: np.ndarray, cl: int, der: int, so: int):
    """coef_imp(5), coef_exp(7) for one grid point.

    class_derivatives_coefficient.f90:32-191.  h0(7) are offsets
    grid(i)-grid(stencil); cl 1=edge, 2=near edge, 3=interior; der 1|2.
    """
    n1, n2 = _DER_STENCIL[cl][so]
    nt = n1 + n2
    h = np.zeros(10)
    if cl == 1:  # :78-79
        h[0:n1] = h0[1:n1 + 1]
        h[n1:nt] = h0[0:n2]
    elif cl == 2:  # :102-103
        h[0] = h0[0]
        if n1 >= 2:
            h[1:n1] = h0[2:n1 + 1]
        h[n1:nt] = h0[0:n2]
    else:  # :122-125 (the second assignment can overwrite h(2))
        i = (4 - n1) // 2
        h[0:2] = h0[i:2 + i]
        lo = 3 - i  # Fortran index
        if n1 >= lo:
            h[lo - 1:n1] = h0[3:n1 + 1 + i]
        i = (5 - n2) // 2
        h[n1:nt] = h0[i:n2 + i]
    a = _taylor_matrix(h, n1, nt, der)
    r = np.zeros(nt)
    r[der] = 1.0
    r = _taylor_solve(a, r)  # dgesv, :159
    coef_imp = np.zeros(5)
    coef_exp = np.zeros(7)
    # export :167-182
    j = 0
    k = (4 - n1) // 2 if cl == 3 else 0
    l = n1 + 1 if n1 + k >= cl else n1
    for i in range(1, l + 1):
        if i != cl - k:
            coef_imp[i + k - 1] = r[j]
            j += 1
    coef_imp[cl - 1] = 1.0
    s = (5 - n2) // 2 if cl == 3 else 0
    coef_exp[s:n2 + s] = r[n1:nt]
    return coef_imp, coef_exp


def cshift(a: np.ndarray, shift: int) -> np.ndarray:
    """Fortran CSHIFT(a, shift)(i) = a(1+MODULO(i-1+shift, n))."""
    n = a.shape[0]
    return a[(np.arange(n) + shift) % n]


def der_coeffs_global(grid: np.ndarray, der: int, so: int):
    """dl(n,5), dr(n,7) for a whole line (class_derivatives.f90:330-377)."""
    n = grid.shape[0]
    dl = np.zeros((n, 5))
    dr = np.zeros((n, 7))
    h = grid[0] - grid[0:7]
    ci, ce = der_coeffs_generic(h, 1, der, so)
    dl[0] = cshift(ci, -2)
    dr[0] = -ce
    h = grid[1] - grid[0:7]
    ci, ce = der_coeffs_generic(h, 2, der, so)
    dl[1] = cshift(ci, -1)
    dr[1] = -ce
    h = np.zeros(7)
    for i in range(2, n - 2):
        h[0:5] = grid[i] - grid[i - 2:i + 3]
        ci, ce = der_coeffs_generic(h, 3, der, so)
        dl[i] = ci
        dr[i] = -ce
    h = grid[n - 2] - grid[n - 1:n - 8:-1] if n > 7 else grid[n - 2] - grid[::-1][:7]
    ci, ce = der_coeffs_generic(h, 2, der, so)
    dl[n - 2] = cshift(ci[::-1], 1)
    dr[n - 2] = -ce[::-1]
    h = grid[n - 1] - grid[n - 1:n - 8:-1] if n > 7 else grid[n - 1] - grid[::-1][:7]
    ci, ce = der_coeffs_generic(h, 1, der, so)
    dl[n - 1] = cshift(ci[::-1], 2)
    dr[n - 1] = -ce[::-1]
    if der == 2:
        dr = -dr
    return dl, dr


def pentalu(dl: np.ndarray) -> np.ndarray:
    """LU of the quasi-pentadiagonal matrix, in the reference's packed form.

    class_derivatives.f90:379-423 with the argument aliasing of :308-309:
    p=dl(:,1), f=dl(:,2), c=dl(:,3), d=dl(:,4), e=dl(:,5),
    alpha=dl(1,1) (= a(1,4)) and beta=dl(n,5) (= a(n,n-3), divided by c(n-3)).
    """
    dl = dl.copy()
    n = dl.shape[0]
    p, f, c, d, e = (dl[:, m] for m in range(5))  # views
    alpha = dl[0, 0]
    f[1] = f[1] / c[0]
    c[1] = c[1] - d[0] * f[1]
    d[1] = d[1] - e[0] * f[1]
    e[1] = e[1] - alpha * f[1]
    i = 2
    p[i] = p[i] / c[i - 2]
    f[i] = (f[i] - p[i] * d[i - 2]) / c[i - 1]
    c[i] = c[i] - d[i - 1] * f[i] - e[i - 2] * p[i]
    d[i] = d[i] - e[i - 1] * f[i] - alpha * p[i]
    for i in range(3, n - 1):
        p[i] = p[i] / c[i - 2]
        f[i] = (f[i] - p[i] * d[i - 2]) / c[i - 1]
        c[i] = c[i] - e[i - 2] * p[i] - d[i - 1] * f[i]
        d[i] = d[i] - e[i - 1] * f[i]
    beta = dl[n - 1, 4] / c[n - 4]
    dl[n - 1, 4] = beta
    p[n - 1] = p[n - 1] - beta * d[n - 4]
    f[n - 1] = f[n - 1] - beta * e[n - 4]
    p[n - 1] = p[n - 1] / c[n - 3]
    f[n - 1] = (f[n - 1] - p[n - 1] * d[n - 3]) / c[n - 2]
    c[n - 1] = c[n - 1] - e[n - 3] * p[n - 1] - d[n - 2] * f[n - 1]
    return dl


def pentares(lu: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Solve with the packed LU (class_derivatives.f90:425-466).

    b has shape (n,) or (n, L): L independent lines solved together.
    """
    n = lu.shape[0]
    p, f, c, d, e = (lu[:, m] for m in range(5))
    alpha = lu[0, 0]
    beta = lu[n - 1, 4]
    x = np.array(b, dtype=np.result_type(b, np.float64), copy=True)
    x[0] = b[0]
    x[1] = b[1] - x[0] * f[1]
    for i in range(2, n - 1):
        x[i] = b[i] - p[i] * x[i - 2] - f[i] * x[i - 1]
    i = n - 1
    x[i] = b[i] - p[i] * x[i - 2] - f[i] * x[i - 1] - beta * x[i - 3]
    x[n - 1] = x[n - 1] / c[n - 1]
    x[n - 2] = (x[n - 2] - d[n - 2] * x[n - 1]) / c[n - 2]
    for i in range(n - 3, 0, -1):
        x[i] = (x[i] - e[i] * x[i + 2] - d[i] * x[i + 1]) / c[i]
    i = 0
    x[i] = (x[i] - e[i] * x[i + 2] - d[i] * x[i + 1] - alpha * x[i + 3]) / c[i]
    return x


def derx_rhs(dr: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Explicit stencil RHS (class_derivatives.f90:262-295); f is (n,) or (n, L).

    Sums run left to right over the stencil like the reference's explicit
    expressions (:277-293).
    """
    n = dr.shape[0]
    f = np.asarray(f)
    f = f.astype(np.result_type(f, np.float64), copy=False)
    ext = (slice(None),) + (None,) * (f.ndim - 1)
    rhs = np.empty_like(f)
    for row in (0, 1):
        acc = dr[row, 0] * f[0]
        for m in range(1, 7):
            acc = acc + dr[row, m] * f[m]
        rhs[row] = acc
    if n > 4:
        mid = dr[2:n - 2, 0][ext] * f[0:n - 4]
        for m in range(1, 5):
            mid = mid + dr[2:n - 2, m][ext] * f[m:n - 4 + m]
        rhs[2:n - 2] = mid
    for row in (n - 2, n - 1):
        acc = dr[row, 0] * f[n - 7]
        for m in range(1, 7):
            acc = acc + dr[row, m] * f[n - 7 + m]
        rhs[row] = acc
    return rhs


class DerCoeffs:
    """type(derivatives_coefficients) (class_derivatives.f90:30-45,111-184)."""

    def __init__(self, grid, solver: bool, so: int = 6):
        grid = np.asarray(grid, dtype=np.float64)
        self.n = grid.shape[0]
        self.so = so
        self.solver = solver
        dl, self.dr = der_coeffs_global(grid, 1, so)
        self.dl = pentalu(dl)
        ddl, self.ddr = der_coeffs_global(grid, 2, so)
        self.ddl = pentalu(ddl)
        if solver:
            dl, self.dr_s = der_coeffs_global(grid[1:-1], 1, so)
            self.dl_s = pentalu(dl)
            ddl, self.ddr_s = der_coeffs_global(grid[1:-1], 2, so)
            self.ddl_s = pentalu(ddl)

    def tables(self, order: int, sub: bool):
        if order == 1:
            return (self.dl_s, self.dr_s) if sub else (self.dl, self.dr)
        return (self.ddl_s, self.ddr_s) if sub else (self.ddl, self.ddr)

    # per-line entry points (class_derivatives.f90:186-260), vectorised over lines
    def der(self, x):
        lu, dr = self.tables(1, False)
        return pentares(lu, derx_rhs(dr, x))

    def dder(self, x):
        lu, dr = self.tables(2, False)
        return pentares(lu, derx_rhs(dr, x))

    def der_s(self, x, out=None):
        lu, dr = self.tables(1, True)
        out = np.zeros_like(x) if out is None else out
        out[1:-1] = pentares(lu, derx_rhs(dr, x[1:-1]))
        return out

    def dder_s(self, x, out=None):
        lu, dr = self.tables(2, True)
        out = np.zeros_like(x) if out is None else out
        out[1:-1] = pentares(lu, derx_rhs(dr, x[1:-1]))
        return out


def derivative_3d(dc: DerCoeffs, x: np.ndarray, axis: int, order: int) -> np.ndarray:
    """derx/dderx/dery/ddery/derz/dderz (class_field.f90:629-978).

    Plain mode: every line (:673-684).  Solver mode: lines in boundary planes
    use the _s coefficients on points 2..n-1, interior lines use the full
    coefficients (:648-671).  Points the reference leaves undefined (line ends
    of boundary-plane lines, and the 4 never-written edge lines) are 0 here.
    """
    x = np.asarray(x)
    x = x.astype(np.result_type(x, np.float64), copy=False)
    xl = np.moveaxis(x, axis, 0)  # (n, a, b)
    n, na, nb = xl.shape
    out = np.zeros_like(xl)
    lu, dr = dc.tables(order, False)
    if not dc.solver:
        lines = xl.reshape(n, na * nb)
        out[:] = pentares(lu, derx_rhs(dr, lines)).reshape(n, na, nb)
    else:
        # interior lines: both transverse indices interior
        inner = xl[:, 1:na - 1, 1:nb - 1].reshape(n, -1)
        out[:, 1:na - 1, 1:nb - 1] = pentares(lu, derx_rhs(dr, inner)).reshape(
            n, na - 2, nb - 2)
        lus, drs = dc.tables(order, True)
        mask = np.zeros((na, nb), dtype=bool)
        mask[0, 1:nb - 1] = mask[na - 1, 1:nb - 1] = True
        mask[1:na - 1, 0] = mask[1:na - 1, nb - 1] = True
        bl = xl[:, mask]  # (n, L)
        res = np.zeros_like(bl)
        res[1:n - 1] = pentares(lus, derx_rhs(drs, bl[1:n - 1]))
        out[:, mask] = res
    out = np.moveaxis(out, 0, axis)
    return np.asfortranarray(out)


# ----------------------------------------------------------------------------
# Solver coefficients (src/class_solver/class_solver_coefficient_3d.f90,
#                      src/class_solver/class_solver_3d.f90:380-656)
# ----------------------------------------------------------------------------


def _solver_stencil(cl: int, so: int, bct: int):
    """(n1, n2) table of class_solver_coefficient_3d.f90:57-168."""
    if cl == 1:
        return {8: (3, 6), 6: (1, 6), 4: (1, 4), 2: (0, 3)}[so]
    if cl == 4:
        return {8: (0, 7), 6: (0, 7), 4: (0, 5), 2: (0, 3)}[so]
    if cl == 2:
        if so == 8:
            return {1: (1, 5), 2: (1, 5), 3: (2, 7)}[bct]
        if so in (6, 4):
            return {1: (1, 5), 2: (1, 5), 3: (1, 3)}[bct]
        return (0, 3)
    return {8: (4, 5), 6: (2, 5), 4: (2, 3), 2: (0, 3)}[so]


def _fortran_div(a: int, b: int) -> int:
    """Fortran integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def solver_coeffs_generic(h0: np.ndarray, cl: int, der: int, bct: int, so: int):
    """coef_imp(5), coef_exp(8) with the boundary-type column.

    class_solver_coefficient_3d.f90:30-245.
    """
    n1, n2 = _solver_stencil(cl, so, bct)
    n3 = n1 + n2
    h = np.zeros(10)
    if cl == 1:  # :76-77
        h[0:n1] = h0[2:n1 + 2]
        h[n1:n3] = h0[0:n2]
    elif cl == 4:  # :96
        h[0:n3] = h0[0:n3]
    elif cl == 2:  # :142-144
        h[0] = h0[1]
        if n1 >= 2:
            h[1:n1] = h0[3:n1 + 2]
        i = max(0, _fortran_div(5 - n2, 2))
        h[n1:n3] = h0[i:n2 + i]
    else:  # :163-166
        i = (4 - n1) // 2
        h[0:2] = h0[i:2 + i]
        lo = 3 - i
        if n1 >= lo:
            h[lo - 1:n1] = h0[3:n1 + 1 + i]
        i = (5 - n2) // 2
        h[n1:n3] = h0[i:n2 + i]
    a = _taylor_matrix(h, n1, n3, der)
    # boundary column (:196-203)
    if (cl == 2 and n2 > 4) or cl == 1:
        c = n1
        a[0:bct - 1, c] = 0.0
        a[bct - 1, c] = -1.0 if bct == 3 else 1.0
        for i in range(bct, n3):  # Fortran i = bct+1..n3
            a[i, c] = a[i - 1, c] * h[c] / (i + 1 - bct)
    r = np.zeros(n3)
    r[der] = 1.0
    r = _taylor_solve(a, r)  # dgesv, :212
    coef_imp = np.zeros(5)
    coef_exp = np.zeros(8)
    j = 0
    k = (4 - n1) // 2 if cl == 3 else 0
    l = n1 + 1 if n1 + k >= cl else n1
    for i in range(1, l + 1):
        if i != cl - k:
            coef_imp[i + k - 1] = r[j]
            j += 1
    if cl <= 5:
        coef_imp[cl - 1] = 1.0
    s = 0
    if cl == 3:
        s = _fortran_div(5 - n2, 2)
    if cl == 2:
        s = max(0, _fortran_div(5 - n2, 2))
    coef_exp[s:n2 + s] = r[n1:n3]
    return coef_imp, coef_exp


def dder_coeffs(grid: np.ndarray, bct, so: int):
    """ddl(n,5), ddr(n,8) on interior points (class_solver_3d.f90:568-619).

    grid has n+2 points (reference grid(0:n+1)).
    """
    g = np.asarray(grid, dtype=np.float64)
    n = g.shape[0] - 2
    der = 2
    ddl = np.zeros((n, 5))
    ddr = np.zeros((n, 8))
    G = lambda i: g[i]  # reference index i in 0..n+1  # noqa: E731
    h = G(1) - g[0:7]
    ci, ce = solver_coeffs_generic(h, 1, der, bct[0], so)
    ddl[0] = cshift(ci, -2)
    ddr[0] = -ce
    if bct[0] == 2:
        ddr[0, 0] = -ddr[0, 0]
    h = G(2) - g[0:7]
    ci, ce = solver_coeffs_generic(h, 2, der, bct[0], so)
    ddl[1] = cshift(ci, -1)
    ddr[1] = -ce
    if bct[0] == 2:
        ddr[1, 0] = -ddr[1, 0]
    h = np.zeros(7)
    for i in range(3, n - 1):  # Fortran i = 3..n-2 on grid(0:n+1)
        h[0:5] = G(i) - g[i - 2:i + 3]
        ci, ce = solver_coeffs_generic(h, 3, der, 0, so)
        ddl[i - 1] = ci
        ddr[i - 1] = -ce
    rev = g[n + 1:n - 6:-1] if n - 6 >= 0 else g[::-1][:7]
    h = G(n - 1) - rev
    ci, ce = solver_coeffs_generic(h, 2, der, bct[1], so)
    ddl[n - 2] = cshift(ci[::-1], 1)
    ddr[n - 2] = -ce[::-1]
    if bct[1] == 2:
        ddr[n - 2, 7] = -ddr[n - 2, 7]
    h = G(n) - rev
    ci, ce = solver_coeffs_generic(h, 1, der, bct[1], so)
    ddl[n - 1] = cshift(ci[::-1], 2)
    ddr[n - 1] = -cshift(ce[::-1], 1)
    if bct[1] == 2:
        ddr[n - 1, 6] = -ddr[n - 1, 6]
    ddr = -ddr  # der == 2, :617
    return ddl, ddr


def solver_lhs_rhs(ddl: np.ndarray, ddr: np.ndarray):
    """Dense lhs(n,n), rhs(n,n) (class_solver_3d.f90:484-521)."""
    n = ddl.shape[0]
    lhs = np.zeros((n, n))
    rhs = np.zeros((n, n))
    lhs[0, 0:4] = [ddl[0, 2], ddl[0, 3], ddl[0, 4], ddl[0, 0]]
    lhs[1, 0:4] = [ddl[1, 1], ddl[1, 2], ddl[1, 3], ddl[1, 4]]
    for i in range(2, n - 2):
        lhs[i, i - 2:i + 3] = ddl[i, 0:5]
    lhs[n - 2, n - 4:n] = ddl[n - 2, 0:4]
    lhs[n - 1, n - 4] = ddl[n - 1, 4]
    lhs[n - 1, n - 3:n] = ddl[n - 1, 0:3]
    rhs[0, 0:6] = ddr[0, 1:7]
    rhs[1, 0:7] = ddr[1, 1:8]
    for i in range(2, n - 2):
        rhs[i, i - 2:i + 3] = ddr[i, 0:5]
    rhs[n - 2, n - 7:n] = ddr[n - 2, 0:7]
    rhs[n - 1, n - 6:n] = ddr[n - 1, 0:6]
    return lhs, rhs


def solver_matrix_init(ddl: np.ndarray, ddr: np.ndarray):
    """eigen, vect, inv_vect, inv_lhs (class_solver_3d.f90:468-566)."""
    lhs, rhs = solver_lhs_rhs(ddl, ddr)
    inv_lhs = np.linalg.inv(lhs)  # dgetrf + dgetri
    temp = inv_lhs @ rhs
    w, v = np.linalg.eig(temp)  # dgeev('N','V')
    if np.max(np.abs(w.imag)) > 1e-13:
        raise ValueError("complex eigenvalues in solver matrix "
                         "(class_solver_3d.f90:537-540): max |Im| = %g"
                         % np.max(np.abs(w.imag)))
    wr = np.ascontiguousarray(w.real)
    vr = np.asfortranarray(v.real)
    inv_v = np.asfortranarray(np.linalg.inv(vr))
    return wr, vr, inv_v, np.asfortranarray(inv_lhs)


NULL_EIGEN_RTOL = 1e-13


def snap_null_eigenvalues(eigen: np.ndarray) -> np.ndarray:
    """Eigenvalues below NULL_EIGEN_RTOL max|lambda| set to exactly 0.

    Deliberate deviation from the reference: the null mode of an axis with
    Neumann conditions at both ends (the constant vector) comes out of dgeev
    as rounding noise, and the solve's null-mode guard uses an absolute
    threshold of 1e-11 (class_solver_3d.f90:207-209) that such noise exceeds
    on strongly stretched grids (1.05e-11 on a 256-point tanh-2.5 axis with
    max|lambda| = 1.5e7); the mode is then divided by the noise and the
    pressure's constant is arbitrary.  With the exact zero the guard acts on
    every grid; where the reference's threshold already catches the mode the
    solve is unchanged."""
    e = np.array(eigen, dtype=np.float64, copy=True)
    e[np.abs(e) <= NULL_EIGEN_RTOL * np.abs(e).max()] = 0.0
    return e


def extrapolation_coeffs(grid: np.ndarray, bct, so: int) -> np.ndarray:
    """neum(2,7) one-sided first-derivative weights (class_solver_3d.f90:622-656).

    The reference passes the literal der=1 for both ends (:647,:653).
    """
    g = np.asarray(grid, dtype=np.float64)
    n = g.shape[0] - 2
    neum = np.zeros((2, 7))
    h = g[0] - g[0:7]
    _, ce = solver_coeffs_generic(h, 4, 1, bct[0], so)
    neum[0] = -ce[0:7]
    rev = g[n + 1:n - 6:-1] if n - 6 >= 0 else g[::-1][:7]
    h = g[n + 1] - rev
    _, ce = solver_coeffs_generic(h, 4, 1, bct[1], so)
    neum[1] = -ce[6::-1]
    return neum


class SolverCoeffs1D:
    """type(solver_coeffs) for one direction (class_solver_3d.f90:28-44,413-466)."""

    def __init__(self, grid, bct, so: int = 6):
        grid = np.asarray(grid, dtype=np.float64)
        self.nall = grid.shape[0]
        self.n = self.nall - 2
        self.so = so
        self.bc = (int(bct[0]), int(bct[1]))
        self.ddl, self.ddr = dder_coeffs(grid, self.bc, so)
        (self.eigen, self.vect, self.inv_vect,
         self.inv_lhs) = solver_matrix_init(self.ddl, self.ddr)
        self.eigen = snap_null_eigenvalues(self.eigen)
        self.neum = extrapolation_coeffs(grid, self.bc, so)
        self.lhs, self.rhs = solver_lhs_rhs(self.ddl, self.ddr)

    def operator(self) -> np.ndarray:
        """Dense A = inv(lhs) rhs (the 1-D operator the diagonalisation inverts)."""
        return self.inv_lhs @ self.rhs


class SolverCoeffs3D:
    """type(solver_coeffs_3d) (class_solver_3d.f90:47-50,380-411).

    As in the reference, the y and z eigenvector matrices are stored
    transposed (:403-409) so that dgy/dgz are right-multiplications.
    """

    def __init__(self, gx, gy, gz, bct, so: int = 6):
        self.cx = SolverCoeffs1D(gx, bct[0:2], so)
        self.cy = SolverCoeffs1D(gy, bct[2:4], so)
        self.cz = SolverCoeffs1D(gz, bct[4:6], so)
        self.vy_t = np.asfortranarray(self.cy.vect.T)
        self.ivy_t = np.asfortranarray(self.cy.inv_vect.T)
        self.vz_t = np.asfortranarray(self.cz.vect.T)
        self.ivz_t = np.asfortranarray(self.cz.inv_vect.T)


def _dgx(a, b):
    """dgemm('n','n',n1,n2*n3,n1) (class_solver_3d.f90:307-323)."""
    n1, n2, n3 = b.shape
    c = a @ b.reshape(n1, n2 * n3, order="F")
    return np.asfortranarray(c.reshape(n1, n2, n3, order="F"))


def _dgy(a, b):
    """per-k dgemm b(:,:,k)·a (class_solver_3d.f90:332-348)."""
    c = np.einsum("ijk,jq->iqk", b, a, optimize=True)
    return np.asfortranarray(c)


def _dgz(a, b):
    """dgemm('n','n',n1*n2,n3,n3) (class_solver_3d.f90:356-378)."""
    n1, n2, n3 = b.shape
    c = b.reshape(n1 * n2, n3, order="F") @ a
    return np.asfortranarray(c.reshape(n1, n2, n3, order="F"))


def fold_bc(sc: SolverCoeffs3D, rhs, bcx, bcy, bcz):
    """BC fold of solve_3d (class_solver_3d.f90:104-175); returns a new array."""
    rhs = np.array(rhs, dtype=np.float64, order="F", copy=True)
    n1, n2, n3 = rhs.shape
    cx, cy, cz = sc.cx, sc.cy, sc.cz
    # x-direction
    c1 = cx.ddr[0, 0] * bcx[:, :, 0]
    c2 = cx.ddr[1, 0] * bcx[:, :, 0]
    c3 = cx.ddr[n1 - 2, 7] * bcx[:, :, 1]
    c4 = cx.ddr[n1 - 1, 6] * bcx[:, :, 1]
    L = cx.inv_lhs
    rhs -= L[:, 0][:, None, None] * c1[None]
    rhs -= L[:, 1][:, None, None] * c2[None]
    rhs -= L[:, n1 - 2][:, None, None] * c3[None]
    rhs -= L[:, n1 - 1][:, None, None] * c4[None]
    # y-direction
    c1 = cy.ddr[0, 0] * bcy[:, :, 0]
    c2 = cy.ddr[1, 0] * bcy[:, :, 0]
    c3 = cy.ddr[n2 - 2, 7] * bcy[:, :, 1]
    c4 = cy.ddr[n2 - 1, 6] * bcy[:, :, 1]
    L = cy.inv_lhs
    rhs -= L[:, 0][None, :, None] * c1[:, None, :]
    rhs -= L[:, 1][None, :, None] * c2[:, None, :]
    rhs -= L[:, n2 - 2][None, :, None] * c3[:, None, :]
    rhs -= L[:, n2 - 1][None, :, None] * c4[:, None, :]
    # z-direction
    c1 = cz.ddr[0, 0] * bcz[:, :, 0]
    c2 = cz.ddr[1, 0] * bcz[:, :, 0]
    c3 = cz.ddr[n3 - 2, 7] * bcz[:, :, 1]
    c4 = cz.ddr[n3 - 1, 6] * bcz[:, :, 1]
    L = cz.inv_lhs
    rhs -= L[:, 0][None, None, :] * c1[:, :, None]
    rhs -= L[:, 1][None, None, :] * c2[:, :, None]
    rhs -= L[:, n3 - 2][None, None, :] * c3[:, :, None]
    rhs -= L[:, n3 - 1][None, None, :] * c4[:, :, None]
    return rhs


def eigen_divide(sc: SolverCoeffs3D, aux, sigma: float):
    """sol = aux/(lx+ly+lz+sigma), 0 at the null mode (class_solver_3d.f90:202-225)."""
    lx = sc.cx.eigen[:, None, None]
    ly = sc.cy.eigen[None, :, None]
    lz = sc.cz.eigen[None, None, :]
    den = lx + ly + lz + sigma
    big = np.maximum(np.maximum(abs(sigma), np.abs(lx)), np.maximum(np.abs(ly), np.abs(lz)))
    null = big < 1e-11
    with np.errstate(divide="ignore", invalid="ignore"):
        sol = np.where(null, 0.0, aux / np.where(null, 1.0, den))
    return np.asfortranarray(sol)


def _extrap(neum_row, vals_in, bc, first: bool):
    """Neumann boundary value (class_solver_3d.f90:658-776)."""
    if first:
        acc = bc.copy()
        for m in range(1, 7):
            acc = acc - neum_row[m] * vals_in[m - 1]
        return acc * (1.0 / neum_row[0])
    acc = bc.copy()
    for m in range(5, -1, -1):
        acc = acc - neum_row[m] * vals_in[m]
    return acc * (1.0 / neum_row[6])


def null_rows(sc: SolverCoeffs3D):
    """Rows of inv(V) of the zero eigenvalue of each axis, or None when some
    axis has none (then sigma = 0 is not singular)."""
    rows = []
    for c in (sc.cx, sc.cy, sc.cz):
        idx = np.flatnonzero(c.eigen == 0.0)
        if len(idx) != 1:
            return None
        rows.append(c.inv_vect[idx[0]])
    return rows


def remove_null_component(sc: SolverCoeffs3D, r):
    """The folded right-hand side minus c0 (the constant vector), c0 =
    w0.r / w0.1 with w0 = w0x (x) w0y (x) w0z the left null vector.

    Deliberate deviation from the reference, equal to it in exact arithmetic:
    the constant is the right null vector of the all-singular operator and
    every other left eigenvector is orthogonal to it, so the modes the
    eigen-divide keeps (class_solver_3d.f90:202-225) do not change.  In
    floating point the incompatible part c0 otherwise leaks into the other
    modes in proportion to the rounding of their left eigenvectors, which
    differs between eigensolvers: on the 192 x 256 x 192 channel the pressure
    moved by 2.5e-8 of its size between two FP64 eigen-decompositions."""
    rows = null_rows(sc)
    if rows is None:
        return r
    wx, wy, wz = rows
    c0 = np.einsum("i,j,k,ijk->", wx, wy, wz, r, optimize=True) / (wx.sum() * wy.sum() * wz.sum())
    return np.asfortranarray(r - c0)


def solve_3d(sc: SolverCoeffs3D, rhs, sigma: float, solf, bcx, bcy, bcz):
    """solve_3d (class_solver_3d.f90:83-299).

    rhs is (n1,n2,n3) interior; solf is the full (n1+2,n2+2,n3+2) field,
    updated in place (edges untouched) and returned.
    """
    n1, n2, n3 = rhs.shape
    r = fold_bc(sc, rhs, bcx, bcy, bcz)
    if sigma == 0.0:
        r = remove_null_component(sc, r)
    aux1 = _dgx(sc.cx.inv_vect, r)
    aux2 = _dgy(sc.ivy_t, aux1)
    aux1 = _dgz(sc.ivz_t, aux2)
    sol = eigen_divide(sc, aux1, sigma)
    aux1 = _dgx(sc.cx.vect, sol)
    aux2 = _dgy(sc.vy_t, aux1)
    sol = _dgz(sc.vz_t, aux2)
    solf[1:n1 + 1, 1:n2 + 1, 1:n3 + 1] = sol
    cx, cy, cz = sc.cx, sc.cy, sc.cz
    # x faces
    if cx.bc[0] == 1:
        solf[0, 1:n2 + 1, 1:n3 + 1] = bcx[:, :, 0]
    elif cx.bc[0] in (2, 3):
        vals = [solf[m, 1:n2 + 1, 1:n3 + 1] for m in range(1, 7)]
        solf[0, 1:n2 + 1, 1:n3 + 1] = _extrap(cx.neum[0], vals, bcx[:, :, 0], True)
    if cx.bc[1] == 1:
        solf[n1 + 1, 1:n2 + 1, 1:n3 + 1] = bcx[:, :, 1]
    elif cx.bc[1] in (2, 3):
        vals = [solf[n1 - 5 + m, 1:n2 + 1, 1:n3 + 1] for m in range(6)]
        solf[n1 + 1, 1:n2 + 1, 1:n3 + 1] = _extrap(cx.neum[1], vals, bcx[:, :, 1], False)
    # y faces
    if cy.bc[0] == 1:
        solf[1:n1 + 1, 0, 1:n3 + 1] = bcy[:, :, 0]
    elif cy.bc[0] in (2, 3):
        vals = [solf[1:n1 + 1, m, 1:n3 + 1] for m in range(1, 7)]
        solf[1:n1 + 1, 0, 1:n3 + 1] = _extrap(cy.neum[0], vals, bcy[:, :, 0], True)
    if cy.bc[1] == 1:
        solf[1:n1 + 1, n2 + 1, 1:n3 + 1] = bcy[:, :, 1]
    elif cy.bc[1] in (2, 3):
        vals = [solf[1:n1 + 1, n2 - 5 + m, 1:n3 + 1] for m in range(6)]
        solf[1:n1 + 1, n2 + 1, 1:n3 + 1] = _extrap(cy.neum[1], vals, bcy[:, :, 1], False)
    # z faces
    if cz.bc[0] == 1:
        solf[1:n1 + 1, 1:n2 + 1, 0] = bcz[:, :, 0]
    elif cz.bc[0] in (2, 3):
        vals = [solf[1:n1 + 1, 1:n2 + 1, m] for m in range(1, 7)]
        solf[1:n1 + 1, 1:n2 + 1, 0] = _extrap(cz.neum[0], vals, bcz[:, :, 0], True)
    if cz.bc[1] == 1:
        solf[1:n1 + 1, 1:n2 + 1, n3 + 1] = bcz[:, :, 1]
    elif cz.bc[1] in (2, 3):
        vals = [solf[1:n1 + 1, 1:n2 + 1, n3 - 5 + m] for m in range(6)]
        solf[1:n1 + 1, 1:n2 + 1, n3 + 1] = _extrap(cz.neum[1], vals, bcz[:, :, 1], False)
    return solf


def solver_3d(sc: SolverCoeffs3D, g, u, bcx, bcy, bcz, sigma: float):
    """Field wrapper solver_3d (class_field.f90:1072-1109): g, u full-size."""
    gr = np.asfortranarray(g[1:-1, 1:-1, 1:-1])
    return solve_3d(sc, gr, sigma, u, bcx, bcy, bcz)


# ----------------------------------------------------------------------------
# Unit-impulse responses (md_influence_matrix_set, class_field.f90:1396-1510)
# ----------------------------------------------------------------------------


def influence_impulses(sc: "SolverCoeffs3D", dcs, sigma: float, face: int, first: int, count: int,
                       exact: bool = False):
    """For each impulse on `face` (0..5 = x1,x2,y1,y2,z1,z2; face points in the
    reference's loop order, i fastest :1431-1432): zero g, u and bc, put a unit
    value in the face's bc (:1435-1459), solve (:1462), take derx/dery/derz
    (:1468-1490) and keep the face-normal derivatives with the sign rule
    (+ at the low face, - at the high face, for m = 1 and m = 2 alike).

    Returns an array (count, 2(n2 n3 + n1 n3 + n1 n2)) of rows
    [bcx.ravel(F) | bcy.ravel(F) | bcz.ravel(F)].  exact=True evaluates the
    solve as the refined discrete solution and the derivatives in 80-bit
    arithmetic (tests only).
    """
    nx, ny, nz = sc.cx.nall, sc.cy.nall, sc.cz.nall
    n1, n2, n3 = nx - 2, ny - 2, nz - 2
    dcx, dcy, dcz = dcs
    rows = []
    for t in range(first, first + count):
        g = np.zeros((nx, ny, nz), order="F")
        u = np.zeros((nx, ny, nz), order="F")
        bcx = np.zeros((n2, n3, 2), order="F")
        bcy = np.zeros((n1, n3, 2), order="F")
        bcz = np.zeros((n1, n2, 2), order="F")
        l, m = face // 2, face % 2
        if l == 0:
            bcx[t % n2, t // n2, m] = 1.0
        elif l == 1:
            bcy[t % n1, t // n1, m] = 1.0
        else:
            bcz[t % n1, t // n1, m] = 1.0
        solver_3d(sc, g, u, bcx, bcy, bcz, sigma)
        if exact:
            ui = solve_exact(sc, g, bcx, bcy, bcz, sigma)
            u = np.array(u, dtype=np.longdouble)
            u[1:-1, 1:-1, 1:-1] = ui
            # faces from the exact interior (Neumann extrapolation)
            for ax, c in enumerate((sc.cx, sc.cy, sc.cz)):
                ul = np.moveaxis(u, ax, 0)
                bc = (bcx, bcy, bcz)[ax]
                inner = (slice(1, -1), slice(1, -1))
                if c.bc[0] != 1:
                    acc = bc[:, :, 0].astype(np.longdouble)
                    for k in range(1, 7):
                        acc = acc - c.neum[0, k] * ul[k][inner]
                    ul[0][inner] = acc * (1.0 / c.neum[0, 0])
                if c.bc[1] != 1:
                    nn = ul.shape[0] - 2
                    acc = bc[:, :, 1].astype(np.longdouble)
                    for k in range(5, -1, -1):
                        acc = acc - c.neum[1, k] * ul[nn - 5 + k][inner]
                    ul[nn + 1][inner] = acc * (1.0 / c.neum[1, 6])
        dx = derivative_3d(dcx, u, 0, 1)
        dy = derivative_3d(dcy, u, 1, 1)
        dz = derivative_3d(dcz, u, 2, 1)
        ox = np.stack([dx[0, 1:-1, 1:-1], -dx[-1, 1:-1, 1:-1]], -1)
        oy = np.stack([dy[1:-1, 0, 1:-1], -dy[1:-1, -1, 1:-1]], -1)
        oz = np.stack([dz[1:-1, 1:-1, 0], -dz[1:-1, 1:-1, -1]], -1)
        rows.append(np.concatenate([ox.ravel(order="F"), oy.ravel(order="F"),
                                    oz.ravel(order="F")]).astype(np.float64))
    return np.array(rows)


# ----------------------------------------------------------------------------
# Exact evaluation helpers (tests only).  Two FP64 implementations of the same
# algorithm differ by their rounding, which for strongly cancelling second
# derivatives and for the diagonalisation solve reaches a few 1e-12; parity is
# therefore measured against the algorithm evaluated in extended precision.
# ----------------------------------------------------------------------------


def derivative_3d_extended(dc: DerCoeffs, x: np.ndarray, axis: int, order: int) -> np.ndarray:
    """derivative_3d with the same FP64 tables, evaluated in 80-bit arithmetic."""
    return derivative_3d(dc, np.asarray(x, dtype=np.longdouble), axis, order)


def _band_apply(mat: np.ndarray, u: np.ndarray, axis: int) -> np.ndarray:
    """mat (n x n, banded) applied along axis, in u's dtype, diagonal by diagonal."""
    n = mat.shape[0]
    ul = np.moveaxis(u, axis, 0)
    out = np.zeros_like(ul)
    m = np.asarray(mat, dtype=u.dtype)
    for off in range(-(n - 1), n):
        d = np.diagonal(m, off)
        if not np.any(d):
            continue
        ext = (slice(None),) + (None,) * (ul.ndim - 1)
        if off >= 0:
            out[0:n - off] += d[ext] * ul[off:n]
        else:
            out[-off:n] += d[ext] * ul[0:n + off]
    return np.moveaxis(out, 0, axis)


def multiplied_residual(sc: "SolverCoeffs3D", u_int, g_int, sigma, bcx, bcy, bcz, dtype=np.longdouble):
    """P (g - L u) with P = lhs_x lhs_y lhs_z, from banded products only.

    P L u = sum_a (prod_{b != a} lhs_b)(rhs_a u + bcol_a bc_a) + sigma P u, the
    operator of class_solver_3d.f90:104-225 multiplied through by P.
    """
    u = np.asarray(u_int, dtype=dtype)
    g = np.asarray(g_int, dtype=dtype)
    cs = (sc.cx, sc.cy, sc.cz)
    bcs = (bcx, bcy, bcz)

    def P(v, skip=None):
        for a in range(3):
            if a != skip:
                v = _band_apply(cs[a].lhs, v, a)
        return v

    res = P(g - sigma * u)
    for a in range(3):
        c = cs[a]
        n = c.n
        t = _band_apply(c.rhs, u, a)
        tl = np.moveaxis(t, a, 0)
        b = np.asarray(bcs[a], dtype=dtype)
        tl[0] += c.ddr[0, 0] * b[:, :, 0]
        tl[1] += c.ddr[1, 0] * b[:, :, 0]
        tl[n - 2] += c.ddr[n - 2, 7] * b[:, :, 1]
        tl[n - 1] += c.ddr[n - 1, 6] * b[:, :, 1]
        res = res - P(t, skip=a)
    return res


def solve_exact(sc: "SolverCoeffs3D", g, bcx, bcy, bcz, sigma, iters: int = 3):
    """Interior solution of the discrete problem to ~1e-15: the FP64 direct
    solve refined with residuals taken in 80-bit arithmetic (the correction
    solves P L d = r' by the FP64 diagonalisation on P^-1 r')."""
    shape = g.shape
    u = np.zeros(shape, order="F")
    solver_3d(sc, g, u, bcx, bcy, bcz, sigma)
    ui = np.array(u[1:-1, 1:-1, 1:-1], dtype=np.longdouble)
    gi = g[1:-1, 1:-1, 1:-1]
    zero = [np.zeros_like(b) for b in (bcx, bcy, bcz)]
    for _ in range(iters):
        rp = multiplied_residual(sc, ui, gi, sigma, bcx, bcy, bcz)
        # r = P^-1 r' (FP64 line solves with the dense lhs)
        r = np.asarray(rp, dtype=np.float64)
        for a, c in enumerate((sc.cx, sc.cy, sc.cz)):
            r = np.moveaxis(np.tensordot(c.inv_lhs, np.moveaxis(r, a, 0), axes=(1, 0)), 0, a)
        d = np.zeros(shape, order="F")
        solve_3d(sc, np.asfortranarray(r), sigma, d, *zero)
        ui = ui + d[1:-1, 1:-1, 1:-1].astype(np.longdouble)
    return ui


# ----------------------------------------------------------------------------
# The discrete operator the diagonalisation inverts (used by the tests to take
# residuals): L u = sum_a inv(lhs_a) (rhs_a u + bcol_a bc_a) + sigma u
# (class_solver_3d.f90:113-124 and :484-532).
# ----------------------------------------------------------------------------


def apply_operator(sc: SolverCoeffs3D, u_int, sigma: float, bcx=None, bcy=None, bcz=None):
    """Apply the 3-D solver operator to an interior field, with optional BCs."""
    u = np.asarray(u_int, dtype=np.float64)
    n1, n2, n3 = u.shape
    out = sigma * u
    for axis, c, bc in ((0, sc.cx, bcx), (1, sc.cy, bcy), (2, sc.cz, bcz)):
        ul = np.moveaxis(u, axis, 0)
        n = ul.shape[0]
        t = np.tensordot(c.rhs, ul, axes=(1, 0))
        if bc is not None:
            b1 = bc[:, :, 0]
            b2 = bc[:, :, 1]
            t[0] += c.ddr[0, 0] * b1
            t[1] += c.ddr[1, 0] * b1
            t[n - 2] += c.ddr[n - 2, 7] * b2
            t[n - 1] += c.ddr[n - 1, 6] * b2
        t = np.tensordot(c.inv_lhs, t, axes=(1, 0))
        out = out + np.moveaxis(t, 0, axis)
    return out

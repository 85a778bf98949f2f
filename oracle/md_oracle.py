"""Multidomain (influence-matrix) Helmholtz/Poisson solve: CPU restatement.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product).

The reference splits the domain into nd(1) x nd(2) x nd(3) subdomains that
share their interface planes, and makes the subdomain solutions C1 across
them through an influence (capacitance) matrix on the interface values:

* interfaces are numbered by md_influence_matrix_param
  (src/class_md/class_md.f90:1450-1560): loop i over x-subdomains, j over y, k
  over z (k fastest); the x-interface above (i,j,k) comes first, then the
  y-, then the z-interface; the unknowns of interface m start at the
  cumulative size of interfaces 1..m-1 (:1621-1630), face points column-major
  over the face's interior;
* column (J, p) of the matrix is the response to a unit Dirichlet value at
  point p of interface J: every subdomain touching J solves with g = 0 and
  that impulse in its boundary data and adds the inward-normal first
  derivatives (+d/dn on its low faces, -d/dn on its high faces) at all its
  interfaces into the rows of those interfaces (md_influence_matrix_set,
  src/class_field/class_field.f90:1396-1510; md_influence_matrix_setvalues,
  class_md.f90:1049-1095, ADD_VALUES);
* the right-hand side comes from the solve with zero interface values:
  -du/dn on the low faces, +du/dn on the high faces, added over the two sides
  (multidomain_solve, class_field.f90:1317-1336; md_vector_setvalues,
  class_md.f90:663-708);
* the interface values x solve M x = rhs (KSPSolve, class_md.f90:379-438) and
  every subdomain solves again with x as its interface data
  (md_vector_getvalues, class_md.f90:445-550; class_field.f90:1365-1369).

Here the interface system is solved densely.  The reference assumes one
subdomain size; this restatement allows a point count per slab (the two
subdomains of a face share their transverse slabs, so faces conform).

Periodic axes (md_influence_matrix_param with periods, class_md.f90:1482-1560):
along a periodic axis the high face of the last subdomain and the low face of
the first are one more interface, numbered where the last subdomain's
interface of that axis would stand in the loop; with one subdomain along
that axis both of its faces carry the same interface.

Singular systems (all outer faces Neumann or periodic, sigma = 0: adding a
constant to every interface value adds it to every subdomain solution, so M
has the constant vector in its null space), the pressure singular method
psm (command_line.f90:27-29, navier_initialization :2119-2121):
* psm = 0: the system as it is;
* psm = 1: the constant is the system's null space: the constant part of
  the right-hand side is removed (md_rhs_nullspace, class_md.f90:347-357)
  and the Krylov solver works orthogonally to the constant (KSPSetNullSpace,
  md_solve_init :318-328), so x has zero mean from a zero start and
  M x = rhs - c 1 for a constant c;
* psm = 2: the first entry of the right-hand side is set to zero
  (md_vector_zero_lastpoint :251-267, applied to the pressure solve only,
  multidomain_solve class_field.f90:1345-1352), and the system is solved as
  it is.

Pinned in tests/test_md.py: a polynomial that the compact schemes
differentiate exactly is reproduced on every subdomain (so any error in the
numbering, the sign rules or the assembly shows); the flux jumps vanish after
the solve; one subdomain reduces to the plain solve.  Periodic layouts: a
periodic trigonometric solution converges at the schemes' order, the flux
jumps vanish across the wrap interface too, and shifting the data by one
slab along the periodic axis shifts the subdomain solutions (a wrong wrap
numbering breaks all three).  psm = 1: M annihilates the constant, the flux
jumps after the solve are one constant (zero for compatible data), x has
zero mean; psm = 2: the first right-hand-side entry is zeroed (checked on a
nonsingular layout against the dense solve of the modified system).
"""
from __future__ import annotations

import numpy as np

import mf_oracle as o


class Layout:
    """nd subdomains per axis; grids[ax][k] = 1-D grid of slab k along ax
    (consecutive slabs share their end points); bct = outer boundary codes
    (x1,x2,y1,y2,z1,z2: 1 Dirichlet, 2 Neumann)."""

    def __init__(self, grids, bct=(1,) * 6, so: int = 6, periods=(False, False, False)):
        self.grids = [list(g) for g in grids]
        self.nd = tuple(len(g) for g in grids)
        self.bct = tuple(bct)
        self.so = so
        self.periods = tuple(bool(p) for p in periods)
        for ax in range(3):
            for k in range(self.nd[ax] - 1):
                if self.grids[ax][k][-1] != self.grids[ax][k + 1][0]:
                    raise ValueError("slabs along axis %d do not share their end points" % ax)
        # subdomains, x fastest
        self.subs = [(a, b, c) for c in range(self.nd[2]) for b in range(self.nd[1])
                     for a in range(self.nd[0])]
        # interfaces (class_md.f90:1525-1559): (axis, low subdomain index)
        self.interfaces = []
        for i in range(self.nd[0]):
            for j in range(self.nd[1]):
                for k in range(self.nd[2]):
                    for ax, q in enumerate((i, j, k)):
                        if q < self.nd[ax] - 1 or self.periods[ax]:
                            self.interfaces.append((ax, (i, j, k)))
        self.iid = {itf: m for m, itf in enumerate(self.interfaces)}
        self.isize = [self.face_size(ax, low) for ax, low in self.interfaces]
        self.ioff = np.concatenate([[0], np.cumsum(self.isize)]).astype(int)
        self.ninf = int(self.ioff[-1])

    def shape(self, s):
        return tuple(len(self.grids[ax][s[ax]]) for ax in range(3))

    def face_size(self, ax, s):
        n = [len(self.grids[q][s[q]]) - 2 for q in range(3)]
        return (n[1] * n[2], n[0] * n[2], n[0] * n[1])[ax]

    def sub_bct(self, s):
        """Boundary codes of subdomain s: interfaces are Dirichlet."""
        b = list(self.bct)
        for f, _ in self.faces(s):
            b[f] = 1
        return tuple(b)

    def faces(self, s):
        """[(face, interface id)] of subdomain s, faces ascending; on a
        periodic axis the first subdomain's low face is the last one's high
        interface."""
        out = []
        for ax in range(3):
            if s[ax] > 0 or self.periods[ax]:
                low = list(s)
                low[ax] = (s[ax] - 1) % self.nd[ax]
                out.append((2 * ax, self.iid[(ax, tuple(low))]))
            if s[ax] < self.nd[ax] - 1 or self.periods[ax]:
                out.append((2 * ax + 1, self.iid[(ax, s)]))
        out.sort()
        return out


def _face_slices(shape):
    nx, ny, nz = shape
    n1, n2, n3 = nx - 2, ny - 2, nz - 2
    sizes = [n2 * n3, n2 * n3, n1 * n3, n1 * n3, n1 * n2, n1 * n2]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    return [slice(offs[f], offs[f + 1]) for f in range(6)]


def _face_normal_derivatives(dcs, u):
    """Row [+d/dn low x | -d/dn high x | ... z] (the influence sign rule) of a
    full field u, faces in the layout of type(boundary_condition)."""
    dx = o.derivative_3d(dcs[0], u, 0, 1)
    dy = o.derivative_3d(dcs[1], u, 1, 1)
    dz = o.derivative_3d(dcs[2], u, 2, 1)
    ox = np.stack([dx[0, 1:-1, 1:-1], -dx[-1, 1:-1, 1:-1]], -1)
    oy = np.stack([dy[1:-1, 0, 1:-1], -dy[1:-1, -1, 1:-1]], -1)
    oz = np.stack([dz[1:-1, 1:-1, 0], -dz[1:-1, 1:-1, -1]], -1)
    return np.concatenate([ox.ravel(order="F"), oy.ravel(order="F"), oz.ravel(order="F")])


class Multidomain:
    def __init__(self, layout: Layout, sigma: float, assemble: bool = True, psm: int = 0):
        """assemble=False skips the dense influence matrix (rhs and
        flux_jumps do not need it; solve does), for layouts too large for a
        dense ninf x ninf matrix.  psm: the pressure singular method (0, 1,
        2; see the module docstring)."""
        if psm not in (0, 1, 2):
            raise ValueError("psm must be 0, 1 or 2")
        self.L = layout
        self.psm = psm
        self.sigma = float(sigma)
        self.sc, self.dcs = {}, {}
        for s in layout.subs:
            gx, gy, gz = (layout.grids[ax][s[ax]] for ax in range(3))
            self.sc[s] = o.SolverCoeffs3D(gx, gy, gz, layout.sub_bct(s), layout.so)
            self.dcs[s] = tuple(o.DerCoeffs(g, solver=True, so=layout.so) for g in (gx, gy, gz))
        self.M = self.influence_matrix() if assemble else None

    def influence_matrix(self):
        """md_influence_matrix_set + setvalues: dense ninf x ninf."""
        L = self.L
        M = np.zeros((L.ninf, L.ninf))
        for s in L.subs:
            sl = _face_slices(L.shape(s))
            faces = L.faces(s)
            for f, J in faces:
                resp = o.influence_impulses(self.sc[s], self.dcs[s], self.sigma, f, 0, L.isize[J])
                cols = slice(L.ioff[J], L.ioff[J + 1])
                for f2, I in faces:
                    M[L.ioff[I]:L.ioff[I + 1], cols] += resp[:, sl[f2]].T
        return M

    def _bc_with(self, s, bc, x):
        """Copies of the subdomain's boundary arrays with its interface faces
        set from x (zero where x is None)."""
        L = self.L
        bcx, bcy, bcz = (np.array(b, dtype=np.float64, order="F") for b in bc)
        arrs = (bcx, bcy, bcz)
        for f, J in L.faces(s):
            ax, side = f // 2, f % 2
            a = arrs[ax]
            vals = 0.0 if x is None else x[L.ioff[J]:L.ioff[J + 1]].reshape(a.shape[:2], order="F")
            a[:, :, side] = vals
        return arrs

    def rhs(self, g, bc):
        """multidomain_solve :1317-1336 with md_vector_setvalues."""
        L = self.L
        r = np.zeros(L.ninf)
        for s in L.subs:
            u = np.zeros(L.shape(s), order="F")
            bcx, bcy, bcz = self._bc_with(s, bc[s], None)
            o.solver_3d(self.sc[s], g[s], u, bcx, bcy, bcz, self.sigma)
            d = -_face_normal_derivatives(self.dcs[s], u)  # -d/dn low, +d/dn high
            sl = _face_slices(L.shape(s))
            for f, J in L.faces(s):
                r[L.ioff[J]:L.ioff[J + 1]] += d[sl[f]]
        return r

    def interface_values(self, r):
        """x from the right-hand side r (KSPSolve with the psm treatment)."""
        n = self.L.ninf
        if n == 0:
            return np.zeros(0)
        if self.psm == 1:
            # (I - P) M x = (I - P) r with x orthogonal to the constant
            # (P: projection on the constant), solved as one consistent
            # least-squares system
            rp = r - r.mean()
            Mp = self.M - self.M.mean(axis=0, keepdims=True)
            A = np.vstack([Mp, np.ones((1, n))])
            return np.linalg.lstsq(A, np.concatenate([rp, [0.0]]), rcond=None)[0]
        r = r.copy()
        if self.psm == 2:
            r[0] = 0.0
        return np.linalg.solve(self.M, r)

    def solve(self, g, bc):
        """Interface values x and the subdomain solutions {s: u}."""
        r = self.rhs(g, bc)
        x = self.interface_values(r)
        u = {}
        for s in self.L.subs:
            us = np.zeros(self.L.shape(s), order="F")
            bcx, bcy, bcz = self._bc_with(s, bc[s], x)
            o.solver_3d(self.sc[s], g[s], us, bcx, bcy, bcz, self.sigma)
            u[s] = us
        return x, u

    def flux_jumps(self, u):
        """For each interface unknown: the sum over its two sides of the
        outward-normal first derivatives (zero for a C1 solution)."""
        L = self.L
        j = np.zeros(L.ninf)
        for s in L.subs:
            d = -_face_normal_derivatives(self.dcs[s], u[s])
            sl = _face_slices(L.shape(s))
            for f, J in L.faces(s):
                j[L.ioff[J]:L.ioff[J + 1]] += d[sl[f]]
        return j


def manufactured(layout: Layout, fun, lap, sigma: float):
    """Per-subdomain g = (lap + sigma) fun and boundary arrays from fun."""
    g, bc = {}, {}
    for s in layout.subs:
        gx, gy, gz = (np.asarray(layout.grids[ax][s[ax]]) for ax in range(3))
        X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
        g[s] = np.asfortranarray(lap(X, Y, Z) + sigma * fun(X, Y, Z))
        U = fun(X, Y, Z)
        bcx = np.asfortranarray(np.stack([U[0, 1:-1, 1:-1], U[-1, 1:-1, 1:-1]], -1))
        bcy = np.asfortranarray(np.stack([U[1:-1, 0, 1:-1], U[1:-1, -1, 1:-1]], -1))
        bcz = np.asfortranarray(np.stack([U[1:-1, 1:-1, 0], U[1:-1, 1:-1, -1]], -1))
        bc[s] = (bcx, bcy, bcz)
    return g, bc
